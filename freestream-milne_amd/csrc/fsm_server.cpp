// fsm_server.cpp -- request loop with continuous batching (fsm_server_* in
// include/fsm.h).  No counterpart in the reference, which runs one event per
// call of run_freestream_milne().
//
// Threads and streams.  fsm_server_submit runs on the caller's thread: it
// finalizes the parameters, checks the shape is supported, reserves the
// request's HBM footprint against the budget, copies the input to pinned
// staging and queues an asynchronous upload on the upload stream.  A worker
// thread owns the launch slots (one non-blocking stream, batch and completion
// event each).  Whenever a slot is free and requests are queued, it takes
// queued requests in arrival order while their estimated device time (from
// the event's plan) stays within max_batch_us and their number within
// max_batch_events, makes the slot's stream wait for their uploads, and
// launches them as one batch (one launch per stage); it retires slots whose
// completion event has fired.  With two or more slots, every slot but the
// last two (at least one) takes only 2-D requests, so a 2-D request never
// waits behind a launch of 3-D events (each of which is about a millisecond of
// device time) and at most two such launches share the GPU with the 2-D
// work: a 3-D event fills the GPU by itself, so a third would only take CU
// slots from 2-D launches, while the second one's workgroups fill the first
// one's tail.  Requests that arrive
// while batches run are therefore admitted and launched next to the running
// work.  Results stay in HBM until fetched or released.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/fsm.h"
#include "fsm_host.h"

namespace {

using Clock = std::chrono::steady_clock;

// Exact-size caching pools: events of a serving mix come in a few shapes, so
// blocks are reused without fragmentation; everything is freed at destroy.
struct DevicePool {
  std::multimap<size_t, void *> free_blocks;
  std::vector<void *> all;
  void *get(size_t bytes) {
    auto it = free_blocks.find(bytes);
    if (it != free_blocks.end()) {
      void *p = it->second;
      free_blocks.erase(it);
      return p;
    }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    all.push_back(p);
    return p;
  }
  void put(size_t bytes, void *p) {
    if (p) free_blocks.emplace(bytes, p);
  }
  void release_all() {
    for (void *p : all) (void)hipFree(p);
    all.clear();
    free_blocks.clear();
  }
};

struct HostPool {
  std::multimap<size_t, void *> free_blocks;
  std::vector<void *> all;
  void *get(size_t bytes) {
    auto it = free_blocks.find(bytes);
    if (it != free_blocks.end()) {
      void *p = it->second;
      free_blocks.erase(it);
      return p;
    }
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
    all.push_back(p);
    return p;
  }
  void put(size_t bytes, void *p) {
    if (p) free_blocks.emplace(bytes, p);
  }
  void release_all() {
    for (void *p : all) (void)hipHostFree(p);
    all.clear();
    free_blocks.clear();
  }
};

enum ReqState { REQ_QUEUED, REQ_RUNNING, REQ_DONE };

struct Request {
  int64_t id = 0;
  fsm_params p{};
  size_t in_bytes = 0, out_bytes = 0, ws_bytes = 0, footprint = 0;
  bool baryon = false;   // input e then n_B; output FSM_NUM_OUTPUTS_BARYON planes
  bool small = true;     // a 2-D event (may use the slot kept for 2-D work)
  float cost_us = 0.0f;  // estimated device time inside a batch
  float *d_in = nullptr;
  double *d_out = nullptr;
  void *ws = nullptr;
  float *h_in = nullptr;
  hipEvent_t uploaded = nullptr;
  ReqState state = REQ_QUEUED;
  int status = FSM_OK;
  Clock::time_point t_submit, t_done;
};

struct Slot {
  fsm_batch *batch = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::vector<int64_t> ids;
  bool busy = false;
  Clock::time_point t_launch;
};

}  // namespace

struct fsm_server {
  fsm_ctx *ctx = nullptr;
  fsm_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::unordered_map<int64_t, Request *> reqs;
  std::deque<int64_t> queued;
  std::deque<int64_t> unreported;
  std::vector<Slot> slots;
  std::vector<hipEvent_t> event_pool;
  hipStream_t upload_stream = nullptr;
  std::mutex pool_mu;  // guards the memory pools (client and worker threads)
  DevicePool dpool;
  HostPool hpool;
  size_t bytes_in_use = 0;
  int64_t next_id = 0, n_admitted = 0, n_completed = 0, n_launches = 0;
  bool stop = false;
  std::thread worker;
};

namespace {

void free_request(fsm_server *s, Request *r) {
  {
    std::lock_guard<std::mutex> g(s->pool_mu);
    s->dpool.put(r->in_bytes, r->d_in);
    s->dpool.put(r->out_bytes, r->d_out);
    s->dpool.put(r->ws_bytes, r->ws);
    s->hpool.put(r->in_bytes, r->h_in);
  }
  if (r->uploaded) s->event_pool.push_back(r->uploaded);  // caller holds s->mu
  delete r;
}

// Worker: launch queued requests into free slots, retire finished slots.
void worker_loop(fsm_server *s) {
  (void)hipSetDevice(fsm_host::ctx_device(s->ctx));
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    // retire finished launches
    for (auto &sl : s->slots) {
      if (!sl.busy) continue;
      hipError_t q = hipEventQuery(sl.done);
      if (q == hipErrorNotReady) continue;
      const Clock::time_point now = Clock::now();
      for (int64_t id : sl.ids) {
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) continue;
        it->second->state = REQ_DONE;
        if (q != hipSuccess) it->second->status = FSM_ERR_DEVICE;
        it->second->t_done = now;
        s->unreported.push_back(id);
        ++s->n_completed;
      }
      sl.ids.clear();
      sl.busy = false;
      s->cv_done.notify_all();
    }
    // compose a launch for the first free slot that has eligible work
    Slot *free_slot = nullptr;
    std::vector<Request *> take;
    for (size_t si = 0; si < s->slots.size() && take.empty() && !s->queued.empty(); ++si) {
      if (s->slots[si].busy) continue;
      const bool small_only = s->slots.size() >= 2 && (int)si < std::max(1, (int)s->slots.size() - 2);
      std::deque<int64_t> skipped;  // stay queued, in order
      float work = 0.0f;
      while (!s->queued.empty()) {
        const int64_t id = s->queued.front();
        auto it = s->reqs.find(id);
        if (it == s->reqs.end() || it->second->state != REQ_QUEUED) {
          s->queued.pop_front();
          continue;
        }
        Request *r = it->second;
        if (small_only && !r->small) {
          skipped.push_back(id);
          s->queued.pop_front();
          continue;
        }
        if ((int)take.size() >= s->cfg.max_batch_events ||
            (!take.empty() && work + r->cost_us > s->cfg.max_batch_us))
          break;  // the launch is full; later requests keep their order
        s->queued.pop_front();
        r->state = REQ_RUNNING;
        work += r->cost_us;
        take.push_back(r);
      }
      s->queued.insert(s->queued.begin(), skipped.begin(), skipped.end());
      if (!take.empty()) free_slot = &s->slots[si];
    }
    if (free_slot) {
      {
        Slot &sl = *free_slot;
        sl.busy = true;
        sl.ids.clear();
        for (Request *r : take) sl.ids.push_back(r->id);
        lk.unlock();
        int st = fsm_batch_clear(sl.batch);
        for (Request *r : take) {
          if (st) break;
          if (hipStreamWaitEvent(sl.stream, r->uploaded, 0) != hipSuccess) st = FSM_ERR_DEVICE;
          int i = fsm_host::batch_add_ws(sl.batch, &r->p, r->d_in, r->baryon ? r->d_in + r->p.DIM : nullptr,
                                         r->d_out, r->ws);
          if (i < 0) st = i;
        }
        if (!st) st = fsm_host::batch_launch_now(sl.batch, sl.stream);
        hipError_t er = hipEventRecord(sl.done, sl.stream);
        lk.lock();
        ++s->n_launches;
        if (st || er != hipSuccess)
          for (Request *r : take) r->status = st ? st : FSM_ERR_DEVICE;
        continue;
      }
    }
    bool any_busy = false;
    for (auto &sl : s->slots) any_busy |= sl.busy;
    if (s->stop && s->queued.empty() && !any_busy) break;
    if (any_busy)
      s->cv_work.wait_for(lk, std::chrono::microseconds(20));
    else
      s->cv_work.wait(lk, [s] { return s->stop || !s->queued.empty(); });
  }
}

}  // namespace

extern "C" {

int fsm_server_create(fsm_ctx *ctx, const fsm_server_config *cfg, fsm_server **out) {
  if (!ctx || !out) return FSM_ERR_ARG;
  *out = nullptr;
  if (hipSetDevice(fsm_host::ctx_device(ctx)) != hipSuccess) return FSM_ERR_DEVICE;
  fsm_server *s = new (std::nothrow) fsm_server();
  if (!s) return FSM_ERR_OOM;
  s->ctx = ctx;
  if (cfg) s->cfg = *cfg;
  if (!s->cfg.hbm_budget_bytes) s->cfg.hbm_budget_bytes = (size_t)16 << 30;
  if (s->cfg.max_batch_events <= 0) s->cfg.max_batch_events = 256;
  if (s->cfg.max_inflight <= 0) s->cfg.max_inflight = 5;
  if (s->cfg.max_requests <= 0) s->cfg.max_requests = 4096;
  // 1 ms of estimated work per launch: a launch costs tens of microseconds of
  // host time (descriptors, uploads, one kernel launch per stage), so small
  // caps bound the request rate; at 3000 requests/s of the mixed 2-D/3-D pool
  // the p99 latency is the same with 0.3 and 1.2 ms caps (about 1.1 ms) and
  // the capacity 18% higher with the larger one
  if (!(s->cfg.max_batch_us > 0.0f)) s->cfg.max_batch_us = 1000.0f;
  int st = FSM_OK;
  if (hipStreamCreateWithFlags(&s->upload_stream, hipStreamNonBlocking) != hipSuccess) st = FSM_ERR_DEVICE;
  s->slots.resize(s->cfg.max_inflight);
  for (auto &sl : s->slots) {
    if (st) break;
    if (fsm_batch_create(ctx, &sl.batch) != FSM_OK) st = FSM_ERR_OOM;
    if (!st && hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess) st = FSM_ERR_DEVICE;
    if (!st && hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) st = FSM_ERR_DEVICE;
    if (!st) fsm_host::batch_set_async_upload(sl.batch, true);
  }
  if (st) {
    fsm_server_destroy(s);
    return st;
  }
  s->worker = std::thread(worker_loop, s);
  *out = s;
  return FSM_OK;
}

void fsm_server_destroy(fsm_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    // queued requests are dropped; launched ones finish
    for (int64_t id : s->queued) {
      auto it = s->reqs.find(id);
      if (it != s->reqs.end() && it->second->state == REQ_QUEUED) it->second->state = REQ_DONE;
    }
    s->queued.clear();
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  (void)hipSetDevice(fsm_host::ctx_device(s->ctx));
  for (auto &sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.batch) fsm_batch_destroy(sl.batch);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.done) (void)hipEventDestroy(sl.done);
  }
  if (s->upload_stream) {
    (void)hipStreamSynchronize(s->upload_stream);
    (void)hipStreamDestroy(s->upload_stream);
  }
  for (auto &kv : s->reqs) {
    if (kv.second->uploaded) (void)hipEventDestroy(kv.second->uploaded);
    delete kv.second;
  }
  for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
  s->dpool.release_all();
  s->hpool.release_all();
  delete s;
}

}  // extern "C"

namespace {

// Admission shared by the host- and device-input forms.  `e_host` (may be
// NULL) is the host copy of the input used by fsm_params_finalize; the input
// itself is copied by `upload` into the request's device buffer on the upload
// stream.
template <typename Upload>
int64_t submit_impl(fsm_server *s, const fsm_params *p, const float *e_host, bool have_nb, size_t n,
                    Upload upload) {
  fsm_params q = *p;
  int st = fsm_params_finalize(&q, e_host, n);
  if (st) return st;
  if (!e_host && n != (size_t)q.DIM) return FSM_ERR_SIZE;
  // BARYON = 1 needs n_B; with BARYON = 0 an n_B is ignored, as the reference ignores charge then
  if (q.BARYON && !have_nb) return FSM_ERR_ARG;
  const bool baryon = q.BARYON && have_nb;
  float cost_us = 0.0f;
  st = fsm_host::plan_check(s->ctx, &q, baryon, &cost_us);
  if (st) return st;
  Request *r = new (std::nothrow) Request();
  if (!r) return FSM_ERR_OOM;
  r->p = q;
  r->baryon = baryon;
  r->small = q.DIM_ETA == 1;
  r->cost_us = cost_us;
  r->in_bytes = (size_t)q.DIM * sizeof(float) * (baryon ? 2 : 1);
  r->out_bytes = (size_t)q.DIM * (baryon ? FSM_NUM_OUTPUTS_BARYON : FSM_NUM_OUTPUTS) * sizeof(double);
  r->ws_bytes = fsm_host::workspace_bytes(q, baryon);
  r->footprint = r->in_bytes + r->out_bytes + r->ws_bytes;
  {
    std::lock_guard<std::mutex> g(s->mu);
    if ((int)s->reqs.size() >= s->cfg.max_requests ||
        s->bytes_in_use + r->footprint > s->cfg.hbm_budget_bytes) {
      delete r;
      return FSM_ERR_BUSY;
    }
    s->bytes_in_use += r->footprint;
    r->id = s->next_id++;
    if (!s->event_pool.empty()) {
      r->uploaded = s->event_pool.back();
      s->event_pool.pop_back();
    }
  }
  bool ok = true;
  {
    std::lock_guard<std::mutex> g(s->pool_mu);
    r->d_in = (float *)s->dpool.get(r->in_bytes);
    r->d_out = (double *)s->dpool.get(r->out_bytes);
    r->ws = s->dpool.get(r->ws_bytes);
    r->h_in = e_host ? (float *)s->hpool.get(r->in_bytes) : nullptr;
    ok = r->d_in && r->d_out && r->ws && (r->h_in || !e_host);
  }
  if (ok && !r->uploaded) ok = hipEventCreateWithFlags(&r->uploaded, hipEventDisableTiming) == hipSuccess;
  if (ok) ok = upload(r, baryon) && hipEventRecord(r->uploaded, s->upload_stream) == hipSuccess;
  std::lock_guard<std::mutex> g(s->mu);
  if (!ok) {
    s->bytes_in_use -= r->footprint;
    free_request(s, r);
    return FSM_ERR_OOM;
  }
  r->t_submit = Clock::now();
  r->state = REQ_QUEUED;
  s->reqs[r->id] = r;
  s->queued.push_back(r->id);
  ++s->n_admitted;
  s->cv_work.notify_one();
  return r->id;
}

}  // namespace

extern "C" {

int64_t fsm_server_submit(fsm_server *s, const fsm_params *p, const float *e_GeV, const float *nB, size_t n) {
  if (!s || !p || !e_GeV) return FSM_ERR_ARG;
  if (p->BARYON && !nB) return FSM_ERR_ARG;
  return submit_impl(s, p, e_GeV, nB != nullptr, n, [&](Request *r, bool baryon) {
    const size_t D = (size_t)r->p.DIM;
    memcpy(r->h_in, e_GeV, D * sizeof(float));
    if (baryon) memcpy(r->h_in + D, nB, D * sizeof(float));
    return hipMemcpyAsync(r->d_in, r->h_in, r->in_bytes, hipMemcpyHostToDevice, s->upload_stream) == hipSuccess;
  });
}

int64_t fsm_server_submit_device(fsm_server *s, const fsm_params *p, const float *d_e_GeV, const float *d_nB,
                                 size_t n, void *stream) {
  if (!s || !p || !d_e_GeV) return FSM_ERR_ARG;
  if (p->BARYON && !d_nB) return FSM_ERR_ARG;
  hipStream_t cs = (hipStream_t)stream;
  // the copies below run on the upload stream after the caller's stream
  hipEvent_t ready = nullptr;
  if (hipEventCreateWithFlags(&ready, hipEventDisableTiming) != hipSuccess) return FSM_ERR_DEVICE;
  if (hipEventRecord(ready, cs) != hipSuccess || hipStreamWaitEvent(s->upload_stream, ready, 0) != hipSuccess) {
    (void)hipEventDestroy(ready);
    return FSM_ERR_DEVICE;
  }
  (void)hipEventDestroy(ready);  // released once the recorded work completes
  // E_DEP_FS needs the reference's serial float sums over the input
  // (src/FreeStream.cpp:389-417): those run on a host copy.
  std::vector<float> host;
  if (p->E_DEP_FS) {
    if (n != (size_t)p->DIM_X * p->DIM_Y * p->DIM_ETA) return FSM_ERR_SIZE;
    host.resize(n);
    if (hipStreamSynchronize(s->upload_stream) != hipSuccess ||
        hipMemcpy(host.data(), d_e_GeV, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      return FSM_ERR_DEVICE;
  }
  return submit_impl(s, p, host.empty() ? nullptr : host.data(), d_nB != nullptr, n, [&](Request *r, bool baryon) {
    const size_t D = (size_t)r->p.DIM;
    bool ok = hipMemcpyAsync(r->d_in, d_e_GeV, D * sizeof(float), hipMemcpyDeviceToDevice, s->upload_stream) ==
              hipSuccess;
    if (ok && baryon)
      ok = hipMemcpyAsync(r->d_in + D, d_nB, D * sizeof(float), hipMemcpyDeviceToDevice, s->upload_stream) ==
           hipSuccess;
    return ok;
  });
}

int64_t fsm_server_poll(fsm_server *s, int timeout_us) {
  if (!s) return FSM_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->unreported.empty() && timeout_us > 0)
    s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), [s] { return !s->unreported.empty(); });
  if (s->unreported.empty()) return -1;
  int64_t id = s->unreported.front();
  s->unreported.pop_front();
  return id;
}

int fsm_server_result(fsm_server *s, int64_t id, double *latency_us, double **d_out) {
  if (!s) return FSM_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return FSM_ERR_ARG;
  Request *r = it->second;
  if (r->state != REQ_DONE) return FSM_ERR_BUSY;
  if (latency_us) latency_us[0] = std::chrono::duration<double, std::micro>(r->t_done - r->t_submit).count();
  if (d_out) *d_out = r->d_out;
  return r->status;
}

int fsm_server_release(fsm_server *s, int64_t id) {
  if (!s) return FSM_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return FSM_ERR_ARG;
  Request *r = it->second;
  if (r->state != REQ_DONE) return FSM_ERR_BUSY;
  s->bytes_in_use -= r->footprint;
  s->reqs.erase(it);
  free_request(s, r);
  return FSM_OK;
}

int fsm_server_fetch(fsm_server *s, int64_t id, double *const out[FSM_NUM_OUTPUTS]) {
  if (!s || !out) return FSM_ERR_ARG;
  double *d_out = nullptr;
  int st = fsm_server_result(s, id, nullptr, &d_out);
  if (st == FSM_ERR_BUSY || st == FSM_ERR_ARG) return st;
  if (st == FSM_OK) {
    size_t D;
    {
      std::lock_guard<std::mutex> g(s->mu);
      D = (size_t)s->reqs[id]->p.DIM;
    }
    (void)hipSetDevice(fsm_host::ctx_device(s->ctx));
    for (int k = 0; k < FSM_NUM_OUTPUTS && st == FSM_OK; ++k)
      if (out[k] && hipMemcpy(out[k], d_out + k * D, D * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        st = FSM_ERR_DEVICE;
  }
  int rs = fsm_server_release(s, id);
  return st ? st : rs;
}

int fsm_server_fetch_baryon(fsm_server *s, int64_t id, double *const out[FSM_NUM_OUTPUTS], double *nB_out,
                            double *const V_out[4]) {
  if (!s || !out) return FSM_ERR_ARG;
  double *d_out = nullptr;
  int st = fsm_server_result(s, id, nullptr, &d_out);
  if (st == FSM_ERR_BUSY || st == FSM_ERR_ARG) return st;
  if (st == FSM_OK) {
    size_t D;
    bool baryon;
    {
      std::lock_guard<std::mutex> g(s->mu);
      D = (size_t)s->reqs[id]->p.DIM;
      baryon = s->reqs[id]->baryon;
    }
    (void)hipSetDevice(fsm_host::ctx_device(s->ctx));
    auto get = [&](double *dst, int plane) {
      if (dst && st == FSM_OK && hipMemcpy(dst, d_out + plane * D, D * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        st = FSM_ERR_DEVICE;
    };
    for (int k = 0; k < FSM_NUM_OUTPUTS; ++k) get(out[k], k);
    if (baryon) {
      get(nB_out, FSM_NUM_OUTPUTS);
      if (V_out)
        for (int c = 0; c < 4; ++c) get(V_out[c], FSM_NUM_OUTPUTS + 1 + c);
    } else if (nB_out || V_out) {
      st = FSM_ERR_ARG;  // not a baryon request
    }
  }
  int rs = fsm_server_release(s, id);
  return st ? st : rs;
}

int fsm_server_cancel(fsm_server *s, int64_t id) {
  if (!s) return FSM_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return FSM_ERR_ARG;
  Request *r = it->second;
  if (r->state != REQ_QUEUED) return FSM_ERR_BUSY;
  auto q = std::find(s->queued.begin(), s->queued.end(), id);
  if (q != s->queued.end()) s->queued.erase(q);
  r->state = REQ_DONE;
  r->status = FSM_ERR_CANCELLED;
  r->t_done = Clock::now();
  s->unreported.push_back(id);
  s->cv_done.notify_all();
  return FSM_OK;
}

int fsm_server_stats(fsm_server *s, int64_t *admitted, int64_t *completed, int64_t *launches,
                     size_t *bytes_in_use) {
  if (!s) return FSM_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  if (admitted) *admitted = s->n_admitted;
  if (completed) *completed = s->n_completed;
  if (launches) *launches = s->n_launches;
  if (bytes_in_use) *bytes_in_use = s->bytes_in_use;
  return FSM_OK;
}

}  // extern "C"
