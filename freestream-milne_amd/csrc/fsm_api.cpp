// fsm_api.cpp -- C ABI of libfsm: context, per-shape planning (stencil and
// spline tables built on the host, uploaded once per event shape) and the
// device-resident and host-buffer entry points declared in include/fsm.h.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/fsm.h"
#include "fsm_host.h"
#include "fsm_internal.h"
#include "fsm_kernels.h"

namespace {

// ----------------------------------------------------------------------------
// Natural-spline derivative operator (host, double).  For unit knot spacing,
// d_i = sum_m G[i][m] z_m gives the slope at knot i of the natural cubic
// spline through z (src/FreeStream.cpp:123,150-153: gsl_interp2d_bicubic,
// whose knot slopes are those of 1-D natural splines along rows and columns).
// Column m of G is the slope vector of the spline through the unit vector e_m.
void natural_slopes_unit(int n, const double *y, double *d) {
  if (n == 1) {
    d[0] = 0.0;
    return;
  }
  std::vector<double> c(n, 0.0);  // c_i = y''_i / 2
  if (n > 2) {
    int m = n - 2;
    std::vector<double> diag(m, 4.0), rhs(m);
    for (int k = 0; k < m; ++k) rhs[k] = 3.0 * (y[k + 2] - 2.0 * y[k + 1] + y[k]);
    for (int k = 1; k < m; ++k) {
      double w = 1.0 / diag[k - 1];
      diag[k] -= w;
      rhs[k] -= w * rhs[k - 1];
    }
    c[m] = rhs[m - 1] / diag[m - 1];
    for (int k = m - 2; k >= 0; --k) c[k + 1] = (rhs[k] - c[k + 2]) / diag[k];
  }
  for (int i = 0; i < n - 1; ++i) d[i] = (y[i + 1] - y[i]) - (c[i + 1] + 2.0 * c[i]) / 3.0;
  double dd = (c[n - 1] - c[n - 2]) / 3.0;
  double b = (y[n - 1] - y[n - 2]) - (c[n - 1] + 2.0 * c[n - 2]) / 3.0;
  d[n - 1] = b + 2.0 * c[n - 2] + 3.0 * dd;
}

// Banded weights gw[i][k], k = 0..2B, acting on z_{i-B+k}; returns the
// largest dropped |G| relative to the largest kept one.
double spline_band(int n, std::vector<float> &gw) {
  const int B = FSM_SPLINE_BAND, W = FSM_SPLINE_W;
  std::vector<double> G((size_t)n * n), e(n), d(n);
  for (int m = 0; m < n; ++m) {
    std::fill(e.begin(), e.end(), 0.0);
    e[m] = 1.0;
    natural_slopes_unit(n, e.data(), d.data());
    for (int i = 0; i < n; ++i) G[(size_t)i * n + m] = d[i];
  }
  gw.assign((size_t)n * W, 0.0f);
  double kept = 0.0, dropped = 0.0;
  for (int i = 0; i < n; ++i)
    for (int m = 0; m < n; ++m) {
      double g = G[(size_t)i * n + m];
      int k = m - i + B;
      if (k >= 0 && k < W) {
        gw[(size_t)i * W + k] = (float)g;
        kept = std::max(kept, fabs(g));
      } else {
        dropped = std::max(dropped, fabs(g));
      }
    }
  return kept > 0 ? dropped / kept : 0.0;
}

inline float hermite(int which, float t) {
  double u = t;
  switch (which) {
    case 0: return (float)((1 + 2 * u) * (1 - u) * (1 - u));
    case 1: return (float)(u * (1 - u) * (1 - u));
    case 2: return (float)(u * u * (3 - 2 * u));
    default: return (float)(u * u * (u - 1));
  }
}

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t ensure(size_t count) {
    if (count <= n && p) return hipSuccess;
    reset();
    hipError_t e = hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) n = count;
    return e;
  }
  hipError_t upload(const std::vector<T> &h) {
    hipError_t e = ensure(h.size());
    if (e != hipSuccess) return e;
    return hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
  }
};

// Everything that depends only on the event shape and time step.
struct Plan {
  fsm_params key{};
  bool valid = false;
  bool is3d = false;
  // 2-D
  Event2D ev2{};
  DevBuf<Stencil2D> st;
  DevBuf<StencilGroup> groups;
  DevBuf<float> gx, gy;
  DevBuf<float4> planes;
  // 3-D
  Event3D ev3{};
  DevBuf<RapSlot> slots;
  DevBuf<float> cs;
  DevBuf<float> src;
  DevBuf<float> Tem;  // eta-major T
};

bool same_shape(const fsm_params &a, const fsm_params &b) { return memcmp(&a, &b, sizeof(a)) == 0; }

}  // namespace

struct fsm_ctx {
  int device = 0;
  int path = FSM_PATH_FP32;
  int timing = 0;
  int max_lds = 65536;
  int num_cus = 256;
  // Timing: one set of 4 events per device call (stage boundaries), pooled
  // so that a run of calls can be summed without synchronising in between.
  std::vector<std::array<hipEvent_t, 4>> pool;
  int used = 0;   // sets recorded since the last collection
  int last = -1;  // set of the most recent call
  float last_ms[3] = {0, 0, 0};
  std::vector<std::shared_ptr<Plan>> plans;
  std::mutex plan_mu;  // the plan cache is shared with the server's worker thread
  // host-API workspace
  DevBuf<float> d_in, d_T;
  DevBuf<double> d_out;
};

namespace {

int hip_status(hipError_t e) {
  if (e == hipSuccess) return FSM_OK;
  if (e == hipErrorOutOfMemory) return FSM_ERR_OOM;
  return FSM_ERR_DEVICE;
}

#define FSM_TRY(expr)                      \
  do {                                     \
    hipError_t e_ = (expr);                \
    if (e_ != hipSuccess) return hip_status(e_); \
  } while (0)

int check_supported(const fsm_params &p) {
  if (p.DIM_X < 4 || p.DIM_Y < 4 || p.DIM_ETA < 1 || p.DIM_RAP < 1 || p.DIM_PHIP < 1) return FSM_ERR_UNSUPPORTED;
  if (!(p.DX > 0) || !(p.DY > 0) || !(p.TAU > 0)) return FSM_ERR_UNSUPPORTED;
  if (p.DIM != p.DIM_X * p.DIM_Y * p.DIM_ETA) return FSM_ERR_ARG;
  if (p.DIM_ETA > 1 && (p.DIM_ETA < 3 || !(p.DETA > 0) || !(p.TAU0 > 0))) return FSM_ERR_UNSUPPORTED;
  return FSM_OK;
}

// ---- 2-D plan ---------------------------------------------------------------
int build_plan2d(fsm_ctx *ctx, const fsm_params &p, Plan &pl) {
  const int nx = p.DIM_X, ny = p.DIM_Y, nphi = p.DIM_PHIP;
  std::vector<float> gxw, gyw;
  double dropx = spline_band(nx, gxw), dropy = spline_band(ny, gyw);
  if (dropx > 1e-9 || dropy > 1e-9) return FSM_ERR_UNSUPPORTED;
  std::vector<Stencil2D> st(nphi);
  const float DTAU = p.DTAU;
  for (int j = 0; j < nphi; ++j) {
    // src/FreeStream.cpp:179,187-188 in float, as the reference evaluates them
    float phip = (float)((double)(float)j * (2.0 * M_PI) / (double)(float)nphi);
    float c = cosf(phip), s = sinf(phip);
    float dxj = c * DTAU, dyj = s * DTAU;
    double sx = -(double)dxj / (double)p.DX, sy = -(double)dyj / (double)p.DY;
    double fx = floor(sx), fy = floor(sy);
    Stencil2D &S = st[j];
    S.ox = (int)fx;
    S.oy = (int)fy;
    float tx = (float)(sx - fx), ty = (float)(sy - fy);
    for (int k = 0; k < 4; ++k) {
      S.hx[k] = hermite(k, tx);
      S.hy[k] = hermite(k, ty);
    }
    S.tx = tx;
    S.ty = ty;
    // hypertrig rows with cosh(0) = 1 (src/LandauMatch.cpp:47-53)
    S.c = c;
    S.s = s;
    S.cc = c * c;
    S.cs = c * s;
    S.ss = s * s;
    S.dxj = dxj;
    S.dyj = dyj;
  }
  // Group contiguous azimuth ranges so that each group's knot box fits the
  // LDS budget (3 workgroups per CU).
  const int budget4 = 3072;  // float4 elements = 48 KiB
  std::vector<StencilGroup> groups;
  int lds_max = 0;
  for (int G = 1; G <= nphi; G *= 2) {
    groups.clear();
    lds_max = 0;
    for (int g = 0; g < G; ++g) {
      int j0 = (int)((long)nphi * g / G), j1 = (int)((long)nphi * (g + 1) / G);
      if (j1 <= j0) continue;
      StencilGroup sg{};
      sg.j0 = j0;
      sg.j1 = j1;
      int oxmin = 1 << 30, oxmax = -(1 << 30), oymin = 1 << 30, oymax = -(1 << 30);
      for (int j = j0; j < j1; ++j) {
        oxmin = std::min(oxmin, st[j].ox);
        oxmax = std::max(oxmax, st[j].ox);
        oymin = std::min(oymin, st[j].oy);
        oymax = std::max(oymax, st[j].oy);
      }
      sg.bx0 = oxmin;
      sg.bx1 = oxmax + 1;
      sg.by0 = oymin;
      sg.by1 = oymax + 1;
      // knot box of the widest tile shape (variant 0) + the group's stencils
      int need = (fsm_stream2d_tile_w(0) + sg.bx1 - sg.bx0) * (fsm_stream2d_tile_h(0) + sg.by1 - sg.by0) +
                 (int)(sizeof(Stencil2D) / sizeof(float4)) * (j1 - j0);
      lds_max = std::max(lds_max, need);
      groups.push_back(sg);
    }
    if (lds_max <= budget4 || G >= nphi) break;
  }
  if (fsm_stream2d_lds_bytes(1, lds_max) > (size_t)ctx->max_lds) return FSM_ERR_UNSUPPORTED;
  // Within a group the order of the samples is free (it only changes the
  // summation order): sort by knot offset so equal offsets are adjacent.
  for (const StencilGroup &sg : groups)
    std::stable_sort(st.begin() + sg.j0, st.begin() + sg.j1, [](const Stencil2D &a, const Stencil2D &b) {
      return a.ox != b.ox ? a.ox < b.ox : a.oy < b.oy;
    });
  FSM_TRY(pl.gx.upload(gxw));
  FSM_TRY(pl.gy.upload(gyw));
  FSM_TRY(pl.st.upload(st));
  FSM_TRY(pl.groups.upload(groups));
  FSM_TRY(pl.planes.ensure((size_t)nx * ny));
  Event2D &ev = pl.ev2;
  memset(&ev, 0, sizeof(ev));
  ev.nx = nx;
  ev.ny = ny;
  ev.ngroups = (int)groups.size();
  ev.lds_floats4 = lds_max;
  ev.out_scale = (float)p.DIM_RAP;
  ev.d_phip = (float)((2.0 * M_PI) / (double)(float)nphi);
  ev.tau = p.TAU;
  // One-event launches pick the widest tile shape that still gives at least
  // two workgroups per CU (the batch path always uses variant 0).
  ev.variant = 2;
  for (int v = 0; v < 3; ++v) {
    int tw = fsm_stream2d_tile_w(v), th = fsm_stream2d_tile_h(v);
    long tiles = (long)((nx + tw - 1) / tw) * ((ny + th - 1) / th);
    if (tiles >= 2L * ctx->num_cus) {
      ev.variant = v;
      break;
    }
  }
  ev.tiles_x = (nx + fsm_stream2d_tile_w(ev.variant) - 1) / fsm_stream2d_tile_w(ev.variant);
  ev.tiles_y = (ny + fsm_stream2d_tile_h(ev.variant) - 1) / fsm_stream2d_tile_h(ev.variant);
  ev.xmin = (float)((-1.0) * ((double)(float)(nx - 1) / 2.0) * (double)p.DX);
  ev.ymin = (float)((-1.0) * ((double)(float)(ny - 1) / 2.0) * (double)p.DY);
  ev.dx = p.DX;
  ev.dy = p.DY;
  ev.planes = pl.planes.p;
  ev.gx = pl.gx.p;
  ev.gy = pl.gy.p;
  ev.st = pl.st.p;
  ev.groups = pl.groups.p;
  ev.t_plane = p.DIM;
  return FSM_OK;
}

// ---- 3-D plan ---------------------------------------------------------------
int build_plan3d(fsm_ctx *ctx, const fsm_params &p, Plan &pl) {
  const int nx = p.DIM_X, ny = p.DIM_Y, neta = p.DIM_ETA, nrap = p.DIM_RAP, nphi = p.DIM_PHIP;
  const float DETA = p.DETA, TAU = p.TAU, TAU0 = p.TAU0, SIGMA = p.SIGMA;
  const float etamin = (float)((-1.0) * ((double)(float)(neta - 1) / 2.0) * (double)DETA);
  const float inv_dx = 1.0f / p.DX, inv_dy = 1.0f / p.DY;
  std::vector<float> cs(2 * (size_t)nphi);
  for (int j = 0; j < nphi; ++j) {
    float phip = (float)((double)(float)j * (2.0 * M_PI) / (double)(float)nphi);
    cs[2 * j] = cosf(phip);
    cs[2 * j + 1] = sinf(phip);
  }
  auto wk = [&](int k) { return (float)(-.9975 + (double)(float)k * (1.995 / (double)(float)(nrap - 1))); };
  auto rap_of = [&](float eta, int k) { return (float)((double)eta + tan((M_PI / 2.0) * (double)wk(k))); };
  // rapidity factor of convertInitialDensity (src/FreeStream.cpp:342-345)
  auto rfac = [&](int ie, int k) {
    float eta = (float)ie * DETA + etamin;
    float rap = rap_of(eta, k);
    float d = eta - rap;
    return (float)((double)(coshf(d) * coshf(d)) *
                   exp((-1.0) * (double)d * (double)d / (2.0 * (double)SIGMA * (double)SIGMA)));
  };
  std::vector<RapSlot> slots((size_t)neta * nrap);
  for (int ie = 0; ie < neta; ++ie) {
    const float eta = (float)ie * DETA + etamin;
    for (int k = 0; k < nrap; ++k) {
      RapSlot S{};
      const float w = wk(k);
      const float rap = rap_of(eta, k);
      // src/FreeStream.cpp:192-194
      const float eta_new = asinhf((TAU / TAU0) * sinhf(eta - rap)) + rap;
      const float D = TAU * coshf(rap - eta_new) - TAU0 * coshf(rap - eta);
      const float ieta_new = (eta_new - etamin) / DETA;
      const float f0 = floorf(ieta_new);
      // src/FreeStream.cpp:250: eta floor index in [1, DIM_ETA-2]; a
      // non-finite D sends every sample out of the transverse range (:213).
      S.valid = (f0 >= 1 && f0 < neta - 1 && std::isfinite(D)) ? 1 : 0;
      if (S.valid) {
        S.eta0 = (int)f0;
        S.feta = ieta_new - f0;
        S.D = D;
        S.rf0 = rfac(S.eta0, k);
        S.rf1 = rfac(S.eta0 + 1, k);
        S.jac = (float)((M_PI / 2.0) / cos((M_PI / 2.0) * (double)w) / cos((M_PI / 2.0) * (double)w) *
                        (1.995 / (double)(float)(nrap - 1)));
        const float ch = coshf(rap - eta), th = tanhf(rap - eta);
        S.inv_ch = 1.0f / ch;
        S.inv_ch2 = 1.0f / (ch * ch);
        S.th_t = (float)((1.0 / (double)TAU) * (double)th);
        S.th_tch = (float)((1.0 / (double)TAU) * (double)th / (double)ch);
        S.th2_t2 = (float)((1.0 / (double)(TAU * TAU)) * (double)th * (double)th);
        int rmin = 1 << 30, rmax = -(1 << 30);
        for (int j = 0; j < nphi; ++j) {
          float sx = -(cs[2 * j] * D) * inv_dx;
          int ox = (int)floorf(sx);
          rmin = std::min(rmin, ox);
          rmax = std::max(rmax, ox);
        }
        S.rmin = rmin;
        S.rmax = rmax;
      }
      slots[(size_t)ie * nrap + k] = S;
    }
  }
  FSM_TRY(pl.slots.upload(slots));
  FSM_TRY(pl.cs.upload(cs));
  FSM_TRY(pl.src.ensure((size_t)p.DIM));
  FSM_TRY(pl.Tem.ensure((size_t)10 * p.DIM));
  Event3D &ev = pl.ev3;
  memset(&ev, 0, sizeof(ev));
  ev.nx = nx;
  ev.ny = ny;
  ev.neta = neta;
  ev.nrap = nrap;
  ev.nphi = nphi;
  ev.tiles = (nx * ny + FSM3D_TILE - 1) / FSM3D_TILE;
  ev.lds_floats = ctx->max_lds / (int)sizeof(float);
  ev.d_phip = (float)((2.0 * M_PI) / (double)(float)nphi);
  ev.inv_dx = inv_dx;
  ev.inv_dy = inv_dy;
  ev.dx = p.DX;
  ev.dy = p.DY;
  ev.xmin = (float)((-1.0) * ((double)(float)(nx - 1) / 2.0) * (double)p.DX);
  ev.ymin = (float)((-1.0) * ((double)(float)(ny - 1) / 2.0) * (double)p.DY);
  ev.src = pl.src.p;
  ev.slots = pl.slots.p;
  ev.cs = pl.cs.p;
  ev.T = pl.Tem.p;
  ev.t_plane = p.DIM;
  {
    double n = sqrt(M_PI / 2.0) * (double)SIGMA * (1.0 + exp(2.0 * (double)SIGMA * (double)SIGMA));
    float nf = (float)n;
    ev.norm_factor = (float)(1.0 / (2.0 * M_PI * (double)nf));
  }
  return FSM_OK;
}

// Plans are cached per parameter set (most recently used first).  A batch
// keeps shared references to the plans of its events, so eviction from the
// cache never frees tables a recorded batch still uses.
int get_plan_shared(fsm_ctx *ctx, const fsm_params &p, std::shared_ptr<Plan> *out) {
  std::lock_guard<std::mutex> lock(ctx->plan_mu);
  for (size_t i = 0; i < ctx->plans.size(); ++i) {
    if (same_shape(ctx->plans[i]->key, p)) {
      std::shared_ptr<Plan> hit = ctx->plans[i];
      ctx->plans.erase(ctx->plans.begin() + i);
      ctx->plans.insert(ctx->plans.begin(), hit);
      *out = hit;
      return FSM_OK;
    }
  }
  int st = check_supported(p);
  if (st) return st;
  std::shared_ptr<Plan> pl(new Plan());
  pl->key = p;
  pl->is3d = p.DIM_ETA > 1;
  st = pl->is3d ? build_plan3d(ctx, p, *pl) : build_plan2d(ctx, p, *pl);
  if (st) return st;
  pl->valid = true;
  ctx->plans.insert(ctx->plans.begin(), pl);
  if (ctx->plans.size() > 64) ctx->plans.pop_back();
  *out = pl;
  return FSM_OK;
}

int get_plan(fsm_ctx *ctx, const fsm_params &p, Plan **out) {
  std::shared_ptr<Plan> sp;
  int st = get_plan_shared(ctx, p, &sp);
  if (st) return st;
  *out = sp.get();
  return FSM_OK;
}

void tmark(fsm_ctx *ctx, int i, hipStream_t s) {
  if (!ctx->timing) return;
  if (i == 0) {
    if (ctx->used == (int)ctx->pool.size()) {
      std::array<hipEvent_t, 4> set{};
      for (auto &e : set)
        if (hipEventCreate(&e) != hipSuccess) {
          ctx->timing = 0;
          return;
        }
      ctx->pool.push_back(set);
    }
    ctx->last = ctx->used++;
  }
  if (ctx->last >= 0) (void)hipEventRecord(ctx->pool[ctx->last][i], s);
}

// Stage A into T with layout chosen by the plan (2-D: reference layout into
// T_out; 3-D: eta-major into the plan's buffer).
int stage_a(fsm_ctx *ctx, Plan &pl, const fsm_params &p, const float *d_e, float *T_ref, hipStream_t s,
            const float **T_used, int *eta_major) {
  if (!pl.is3d) {
    Event2D ev = pl.ev2;
    ev.e_gev = d_e;
    ev.T = T_ref;
    tmark(ctx, 0, s);
    FSM_TRY(fsm_launch_spline2d(ev, s));
    tmark(ctx, 1, s);
    FSM_TRY(fsm_launch_stream2d(ev, s));
    tmark(ctx, 2, s);
    *T_used = T_ref;
    *eta_major = 0;
  } else {
    Event3D ev = pl.ev3;
    ev.e_gev = d_e;
    tmark(ctx, 0, s);
    FSM_TRY(fsm_launch_src3d(ev, s));
    tmark(ctx, 1, s);
    FSM_TRY(fsm_launch_stream3d(ev, s));
    tmark(ctx, 2, s);
    *T_used = pl.Tem.p;
    *eta_major = 1;
  }
  return FSM_OK;
}

EventLandau landau_event(const fsm_params &p, const float *T, int eta_major, int mixed) {
  EventLandau L;
  memset(&L, 0, sizeof(L));
  L.dim = p.DIM;
  L.nxy = p.DIM_X * p.DIM_Y;
  L.neta = p.DIM_ETA;
  L.t_eta_major = eta_major;
  L.eos_type = p.EOS_TYPE;
  L.viscous = p.VISCOUS_MATCHING;
  L.fill_outputs = (p.OUTPUTFORMAT == 2 || p.OUTPUTFORMAT == 3) ? 1 : 0;
  L.mixed = mixed;
  L.tau = p.TAU;
  L.T = T;
  L.t_plane = p.DIM;
  L.out_plane = p.DIM;
  return L;
}

void finish_timing(fsm_ctx *ctx) {
  if (ctx->last < 0) return;
  auto &set = ctx->pool[ctx->last];
  (void)hipEventSynchronize(set[3]);
  for (int i = 0; i < 3; ++i) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, set[i], set[i + 1]) == hipSuccess) ctx->last_ms[i] = ms;
  }
}

int run_device_impl(fsm_ctx *ctx, const fsm_params *p, const float *d_e, double *d_out, float *T_scratch,
                    hipStream_t s) {
  Plan *pl = nullptr;
  int st = get_plan(ctx, *p, &pl);
  if (st) return st;
  const float *T = nullptr;
  int em = 0;
  st = stage_a(ctx, *pl, *p, d_e, T_scratch, s, &T, &em);
  if (st) return st;
  EventLandau L = landau_event(*p, T, em, ctx->path == FSM_PATH_MIXED);
  L.out = d_out;
  FSM_TRY(fsm_launch_landau(L, s));
  tmark(ctx, 3, s);
  return FSM_OK;
}

}  // namespace

extern "C" {

const char *fsm_version(void) { return "fsm-mi355x 0.1"; }

const char *fsm_status_string(int status) {
  switch (status) {
    case FSM_OK: return "ok";
    case FSM_ERR_SIZE: return "input vector size does not match DIM_X*DIM_Y*DIM_ETA";
    case FSM_ERR_IC: return "initial condition not supported on this path";
    case FSM_ERR_FILE: return "file could not be opened";
    case FSM_ERR_OOM: return "out of memory";
    case FSM_ERR_DEVICE: return "HIP device error";
    case FSM_ERR_ARG: return "invalid argument";
    case FSM_ERR_UNSUPPORTED: return "parameter combination not supported";
    case FSM_ERR_BUSY: return "no free slot";
    case FSM_ERR_CANCELLED: return "cancelled";
    default: return "unknown status";
  }
}

int fsm_create(int device, fsm_ctx **out) {
  if (!out) return FSM_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return FSM_ERR_DEVICE;
  FSM_TRY(hipSetDevice(device));
  std::unique_ptr<fsm_ctx> ctx(new fsm_ctx());
  ctx->device = device;
  hipDeviceProp_t prop;
  FSM_TRY(hipGetDeviceProperties(&prop, device));
  // 3-D planes are staged whole when they fit; keep two workgroups per CU.
  int lds = (int)std::min<size_t>(prop.sharedMemPerBlock, (size_t)80 * 1024);
  ctx->max_lds = lds;
  ctx->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  FSM_TRY(fsm_set_lds_limits(lds, lds));
  *out = ctx.release();
  return FSM_OK;
}

void fsm_destroy(fsm_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  for (auto &set : ctx->pool)
    for (auto &e : set) (void)hipEventDestroy(e);
  delete ctx;
}

int fsm_set_path(fsm_ctx *ctx, int path) {
  if (!ctx) return FSM_ERR_ARG;
  if (path != FSM_PATH_FP32 && path != FSM_PATH_MIXED) return FSM_ERR_ARG;
  ctx->path = path;
  return FSM_OK;
}

int fsm_enable_timing(fsm_ctx *ctx, int on) {
  if (!ctx) return FSM_ERR_ARG;
  ctx->timing = on ? 1 : 0;
  return FSM_OK;
}

int fsm_last_kernel_ms(fsm_ctx *ctx, int which, float *ms) {
  if (!ctx || !ms || which < 0 || which > 2) return FSM_ERR_ARG;
  finish_timing(ctx);
  *ms = ctx->last_ms[which];
  return FSM_OK;
}

int fsm_kernel_time_totals(fsm_ctx *ctx, double ms[3], int *ncalls) {
  if (!ctx || !ms) return FSM_ERR_ARG;
  ms[0] = ms[1] = ms[2] = 0.0;
  if (ctx->used > 0) FSM_TRY(hipEventSynchronize(ctx->pool[ctx->used - 1][3]));
  for (int k = 0; k < ctx->used; ++k)
    for (int i = 0; i < 3; ++i) {
      float t = 0.0f;
      if (hipEventElapsedTime(&t, ctx->pool[k][i], ctx->pool[k][i + 1]) == hipSuccess) ms[i] += t;
    }
  if (ncalls) *ncalls = ctx->used;
  ctx->used = 0;
  ctx->last = -1;
  return FSM_OK;
}

int fsm_sync(fsm_ctx *ctx, void *stream) {
  if (!ctx) return FSM_ERR_ARG;
  FSM_TRY(hipSetDevice(ctx->device));
  FSM_TRY(hipStreamSynchronize((hipStream_t)stream));
  return FSM_OK;
}

int fsm_run_device(fsm_ctx *ctx, const fsm_params *p, const float *d_e_GeV, double *d_out, void *stream) {
  if (!ctx || !p || !d_e_GeV || !d_out) return FSM_ERR_ARG;
  FSM_TRY(hipSetDevice(ctx->device));
  float *Tscr = nullptr;
  if (p->DIM_ETA == 1) {
    FSM_TRY(ctx->d_T.ensure((size_t)10 * p->DIM));
    Tscr = ctx->d_T.p;
  }
  return run_device_impl(ctx, p, d_e_GeV, d_out, Tscr, (hipStream_t)stream);
}

int fsm_stream_tmunu_device(fsm_ctx *ctx, const fsm_params *p, const float *d_e_GeV, float *d_T, void *stream) {
  if (!ctx || !p || !d_e_GeV || !d_T) return FSM_ERR_ARG;
  FSM_TRY(hipSetDevice(ctx->device));
  Plan *pl = nullptr;
  int st = get_plan(ctx, *p, &pl);
  if (st) return st;
  hipStream_t s = (hipStream_t)stream;
  const float *T = nullptr;
  int em = 0;
  st = stage_a(ctx, *pl, *p, d_e_GeV, d_T, s, &T, &em);
  if (st) return st;
  if (em) {
    // eta-major [10][neta][nxy] -> reference [10][nxy][neta]
    FSM_TRY(fsm_launch_transpose(T, d_T, p->DIM_ETA, p->DIM_X * p->DIM_Y, 10, p->DIM, p->DIM, s));
  }
  tmark(ctx, 3, s);
  return FSM_OK;
}

int fsm_landau_device(fsm_ctx *ctx, const fsm_params *p, const float *d_T, float *d_e, float *d_u, float *d_p,
                      float *d_Pi, float *d_pi, void *stream) {
  if (!ctx || !p || !d_T) return FSM_ERR_ARG;
  FSM_TRY(hipSetDevice(ctx->device));
  int st = check_supported(*p);
  if (st) return st;
  hipStream_t s = (hipStream_t)stream;
  EventLandau L = landau_event(*p, d_T, 0, ctx->path == FSM_PATH_MIXED);
  L.fe = d_e;
  L.fu = d_u;
  L.fp = d_p;
  L.fPi = d_Pi;
  L.fpi = d_pi;
  tmark(ctx, 0, s);
  tmark(ctx, 1, s);
  tmark(ctx, 2, s);
  FSM_TRY(fsm_launch_landau(L, s));
  tmark(ctx, 3, s);
  return FSM_OK;
}

int fsm_run(fsm_ctx *ctx, fsm_params *p, const float *e_GeV, size_t n, double *const out[FSM_NUM_OUTPUTS],
            float *tau_landau_match) {
  if (!ctx || !p || !e_GeV || !out) return FSM_ERR_ARG;
  for (int k = 0; k < FSM_NUM_OUTPUTS; ++k)
    if (!out[k]) return FSM_ERR_ARG;
  int st = fsm_params_finalize(p, e_GeV, n);
  if (st) return st;
  if (tau_landau_match) *tau_landau_match = p->TAU;
  FSM_TRY(hipSetDevice(ctx->device));
  const size_t D = (size_t)p->DIM;
  FSM_TRY(ctx->d_in.ensure(D));
  FSM_TRY(ctx->d_out.ensure(D * FSM_NUM_OUTPUTS));
  hipStream_t s = nullptr;
  FSM_TRY(hipMemcpyAsync(ctx->d_in.p, e_GeV, D * sizeof(float), hipMemcpyHostToDevice, s));
  st = fsm_run_device(ctx, p, ctx->d_in.p, ctx->d_out.p, s);
  if (st) return st;
  for (int k = 0; k < FSM_NUM_OUTPUTS; ++k)
    FSM_TRY(hipMemcpyAsync(out[k], ctx->d_out.p + k * D, D * sizeof(double), hipMemcpyDeviceToHost, s));
  FSM_TRY(hipStreamSynchronize(s));
  return FSM_OK;
}

int fsm_run_baryon(fsm_ctx *ctx, fsm_params *p, const float *e_GeV, const float *nB_in, size_t n,
                   double *const out[FSM_NUM_OUTPUTS], double *nB_out, double *const V_out[4],
                   float *tau_landau_match) {
  (void)nB_in;
  (void)nB_out;
  (void)V_out;
  if (!p) return FSM_ERR_ARG;
  if (p->BARYON) return FSM_ERR_UNSUPPORTED;
  return fsm_run(ctx, p, e_GeV, n, out, tau_landau_match);
}

}  // extern "C"

// ============================================================================
// Batches
// ============================================================================
namespace {

struct BatchEvent {
  std::shared_ptr<Plan> plan;
  fsm_params p;
  const float *d_e;
  double *d_out;
  char *ws = nullptr;  // caller-provided workspace (event_workspace_bytes), else the batch arena
};

inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Device workspace of one event in a batch: spline planes + T (2-D), or
// source planes + eta-major T (3-D).
size_t event_workspace_bytes(const fsm_params &p) {
  if (p.DIM_ETA == 1)
    return align256((size_t)p.DIM_X * p.DIM_Y * sizeof(float4)) + align256((size_t)10 * p.DIM * sizeof(float));
  return align256((size_t)p.DIM * sizeof(float)) + align256((size_t)10 * p.DIM * sizeof(float));
}

// Host staging for stream-ordered uploads (pinned, grown as needed).
struct PinnedBuf {
  char *p = nullptr;
  size_t n = 0;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  hipError_t ensure(size_t bytes) {
    if (bytes <= n && p) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipHostMalloc((void **)&p, std::max<size_t>(bytes, 256), hipHostMallocDefault);
    if (e == hipSuccess) n = std::max<size_t>(bytes, 256);
    return e;
  }
};

}  // namespace

struct fsm_batch {
  fsm_ctx *ctx = nullptr;
  std::vector<BatchEvent> evs;
  bool dirty = true;
  int use_graph = 1;
  int built_path = -1;  // context path the descriptors were built for
  DevBuf<char> arena;
  DevBuf<Event2D> d_ev2;
  DevBuf<Event3D> d_ev3;
  DevBuf<EventLandau> d_evl;
  DevBuf<int4> d_items;
  int off_sp = 0, n_sp = 0, off_s2 = 0, n_s2 = 0, off_src = 0, n_src = 0, off_s3 = 0, n_s3 = 0, off_l = 0,
      n_l = 0;
  size_t lds2 = 0, lds3 = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // stream-ordered uploads (used by the server): descriptors go through
  // pinned staging with hipMemcpyAsync on the launch stream
  bool async_upload = false;
  PinnedBuf staging;

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~fsm_batch() { drop_graph(); }
};

namespace {

const int kLandauChunk = 1024;

int batch_build(fsm_batch *b, hipStream_t s) {
  b->drop_graph();
  b->built_path = b->ctx->path;
  std::vector<Event2D> e2;
  std::vector<Event3D> e3;
  std::vector<EventLandau> el;
  std::vector<int4> sp, s2, src, s3, ld;
  size_t bytes = 0;
  for (auto &ev : b->evs)
    if (!ev.ws) bytes += event_workspace_bytes(ev.p);
  if (bytes) FSM_TRY(b->arena.ensure(bytes));
  char *arena_cur = b->arena.p;
  b->lds2 = b->lds3 = 0;
  for (auto &ev : b->evs) {
    const fsm_params &p = ev.p;
    const int li = (int)el.size();
    char *cur = ev.ws;
    if (!cur) {
      cur = arena_cur;
      arena_cur += event_workspace_bytes(p);
    }
    if (!ev.plan->is3d) {
      Event2D d = ev.plan->ev2;
      d.e_gev = ev.d_e;
      d.planes = (float4 *)cur;
      cur += align256((size_t)p.DIM_X * p.DIM_Y * sizeof(float4));
      d.T = (float *)cur;
      cur += align256((size_t)10 * p.DIM * sizeof(float));
      d.t_plane = p.DIM;
      const int i = (int)e2.size();
      e2.push_back(d);
      const int stx = (p.DIM_X + FSMSP_TX - 1) / FSMSP_TX, sty = (p.DIM_Y + FSMSP_TY - 1) / FSMSP_TY;
      for (int x = 0; x < stx; ++x)
        for (int y = 0; y < sty; ++y) sp.push_back(make_int4(i, x, y, 0));
      // each event keeps its plan's tile shape (identical bits alone and batched)
      for (int y = 0; y < d.tiles_y; ++y)
        for (int x = 0; x < d.tiles_x; ++x) s2.push_back(make_int4(i, x, y, 0));
      b->lds2 = std::max(b->lds2, fsm_stream2d_lds_bytes(d.variant, d.lds_floats4));
      EventLandau L = landau_event(p, d.T, 0, b->ctx->path == FSM_PATH_MIXED);
      L.out = ev.d_out;
      el.push_back(L);
    } else {
      Event3D d = ev.plan->ev3;
      d.e_gev = ev.d_e;
      d.src = (float *)cur;
      cur += align256((size_t)p.DIM * sizeof(float));
      d.T = (float *)cur;
      cur += align256((size_t)10 * p.DIM * sizeof(float));
      d.t_plane = p.DIM;
      const int i = (int)e3.size();
      e3.push_back(d);
      const int nxy = p.DIM_X * p.DIM_Y;
      for (int x = 0; x < (nxy + 63) / 64; ++x)
        for (int y = 0; y < (p.DIM_ETA + 63) / 64; ++y) src.push_back(make_int4(i, x, y, 0));
      for (int ie = 0; ie < p.DIM_ETA; ++ie)
        for (int t = 0; t < d.tiles; ++t) s3.push_back(make_int4(i, t, ie, 0));
      b->lds3 = std::max(b->lds3, (size_t)d.lds_floats * sizeof(float));
      EventLandau L = landau_event(p, d.T, 1, b->ctx->path == FSM_PATH_MIXED);
      L.out = ev.d_out;
      el.push_back(L);
    }
    for (int c = 0; c < p.DIM; c += kLandauChunk) ld.push_back(make_int4(li, c, std::min(kLandauChunk, p.DIM - c), 0));
  }
  std::vector<int4> items;
  auto put = [&](std::vector<int4> &v, int *off, int *n) {
    *off = (int)items.size();
    *n = (int)v.size();
    items.insert(items.end(), v.begin(), v.end());
  };
  put(sp, &b->off_sp, &b->n_sp);
  put(s2, &b->off_s2, &b->n_s2);
  put(src, &b->off_src, &b->n_src);
  put(s3, &b->off_s3, &b->n_s3);
  put(ld, &b->off_l, &b->n_l);
  if (!b->async_upload) {
    if (!e2.empty()) FSM_TRY(b->d_ev2.upload(e2));
    if (!e3.empty()) FSM_TRY(b->d_ev3.upload(e3));
    FSM_TRY(b->d_evl.upload(el));
    FSM_TRY(b->d_items.upload(items));
  } else {
    // one pinned staging block, four stream-ordered copies
    const size_t n2 = e2.size() * sizeof(Event2D), n3 = e3.size() * sizeof(Event3D),
                 nl = el.size() * sizeof(EventLandau), ni = items.size() * sizeof(int4);
    const size_t o3 = align256(n2), ol = o3 + align256(n3), oi = ol + align256(nl);
    FSM_TRY(b->staging.ensure(oi + ni));
    memcpy(b->staging.p, e2.data(), n2);
    memcpy(b->staging.p + o3, e3.data(), n3);
    memcpy(b->staging.p + ol, el.data(), nl);
    memcpy(b->staging.p + oi, items.data(), ni);
    FSM_TRY(b->d_ev2.ensure(e2.size()));
    FSM_TRY(b->d_ev3.ensure(e3.size()));
    FSM_TRY(b->d_evl.ensure(el.size()));
    FSM_TRY(b->d_items.ensure(items.size()));
    if (n2) FSM_TRY(hipMemcpyAsync(b->d_ev2.p, b->staging.p, n2, hipMemcpyHostToDevice, s));
    if (n3) FSM_TRY(hipMemcpyAsync(b->d_ev3.p, b->staging.p + o3, n3, hipMemcpyHostToDevice, s));
    FSM_TRY(hipMemcpyAsync(b->d_evl.p, b->staging.p + ol, nl, hipMemcpyHostToDevice, s));
    FSM_TRY(hipMemcpyAsync(b->d_items.p, b->staging.p + oi, ni, hipMemcpyHostToDevice, s));
  }
  b->dirty = false;
  return FSM_OK;
}

int batch_launch(fsm_batch *b, hipStream_t s, bool timed) {
  fsm_ctx *ctx = b->ctx;
  const int4 *it = b->d_items.p;
  if (timed) tmark(ctx, 0, s);
  FSM_TRY(fsm_launch_spline2d_batch(b->d_ev2.p, it + b->off_sp, b->n_sp, s));
  FSM_TRY(fsm_launch_src3d_batch(b->d_ev3.p, it + b->off_src, b->n_src, s));
  if (timed) tmark(ctx, 1, s);
  FSM_TRY(fsm_launch_stream2d_batch(b->d_ev2.p, it + b->off_s2, b->n_s2, b->lds2, s));
  FSM_TRY(fsm_launch_stream3d_batch(b->d_ev3.p, it + b->off_s3, b->n_s3, b->lds3, s));
  if (timed) tmark(ctx, 2, s);
  FSM_TRY(fsm_launch_landau_batch(b->d_evl.p, it + b->off_l, b->n_l, s));
  if (timed) tmark(ctx, 3, s);
  return FSM_OK;
}

}  // namespace

extern "C" {

int fsm_batch_create(fsm_ctx *ctx, fsm_batch **out) {
  if (!ctx || !out) return FSM_ERR_ARG;
  fsm_batch *b = new (std::nothrow) fsm_batch();
  if (!b) return FSM_ERR_OOM;
  b->ctx = ctx;
  *out = b;
  return FSM_OK;
}

void fsm_batch_destroy(fsm_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  delete b;
}

int fsm_batch_add(fsm_batch *b, const fsm_params *p, const float *d_e_GeV, double *d_out) {
  if (!b || !p || !d_e_GeV || !d_out) return FSM_ERR_ARG;
  FSM_TRY(hipSetDevice(b->ctx->device));
  BatchEvent ev;
  int st = get_plan_shared(b->ctx, *p, &ev.plan);
  if (st) return st;
  ev.p = *p;
  ev.d_e = d_e_GeV;
  ev.d_out = d_out;
  b->evs.push_back(ev);
  b->dirty = true;
  return (int)b->evs.size() - 1;
}

int fsm_batch_clear(fsm_batch *b) {
  if (!b) return FSM_ERR_ARG;
  b->evs.clear();
  b->dirty = true;
  b->drop_graph();
  return FSM_OK;
}

int fsm_batch_size(const fsm_batch *b) { return b ? (int)b->evs.size() : FSM_ERR_ARG; }

int fsm_batch_set_graph(fsm_batch *b, int use_graph) {
  if (!b) return FSM_ERR_ARG;
  b->use_graph = use_graph ? 1 : 0;
  return FSM_OK;
}

int fsm_batch_run(fsm_batch *b, void *stream) {
  if (!b) return FSM_ERR_ARG;
  if (b->evs.empty()) return FSM_OK;
  FSM_TRY(hipSetDevice(b->ctx->device));
  hipStream_t s = (hipStream_t)stream;
  if (b->dirty || b->built_path != b->ctx->path) {
    int st = batch_build(b, s);
    if (st) return st;
  }
  if (!b->use_graph || s == nullptr) return batch_launch(b, s, b->ctx->timing != 0);
  if (!b->exec) {
    FSM_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int st = batch_launch(b, s, false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (st) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    FSM_TRY(e);
    b->graph = g;
    FSM_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  FSM_TRY(hipGraphLaunch(b->exec, s));
  return FSM_OK;
}

}  // extern "C"

// ---- internal entry points for the server (fsm_host.h) ----------------------
namespace fsm_host {

size_t workspace_bytes(const fsm_params &p) { return event_workspace_bytes(p); }

int batch_add_ws(fsm_batch *b, const fsm_params *p, const float *d_e, double *d_out, void *ws) {
  int i = fsm_batch_add(b, p, d_e, d_out);
  if (i >= 0) b->evs[i].ws = (char *)ws;
  return i;
}

void batch_set_async_upload(fsm_batch *b, bool on) { b->async_upload = on; }

int batch_launch_now(fsm_batch *b, hipStream_t s) {
  if (b->evs.empty()) return FSM_OK;
  if (b->dirty || b->built_path != b->ctx->path) {
    int st = batch_build(b, s);
    if (st) return st;
  }
  return batch_launch(b, s, false);
}

int ctx_device(const fsm_ctx *ctx) { return ctx->device; }

int plan_check(fsm_ctx *ctx, const fsm_params *p) {
  std::shared_ptr<Plan> sp;
  return get_plan_shared(ctx, *p, &sp);
}

}  // namespace fsm_host
