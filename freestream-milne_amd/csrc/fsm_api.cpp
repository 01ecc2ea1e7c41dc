// fsm_api.cpp -- C ABI of libfsm: context, per-shape planning (stencil and
// spline tables built on the host, uploaded once per event shape) and the
// device-resident and host-buffer entry points declared in include/fsm.h.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <sys/stat.h>

#include <algorithm>
#include <array>
#include <map>
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
  int rapidity_rule = FSM_RAPIDITY_TAN;  // 3-D plans: the rule their rapidity tables follow
  // 2-D
  Event2D ev2{};
  DevBuf<Stencil2D> st;
  DevBuf<StencilGroup> groups;
  DevBuf<float> gx, gy;
  DevBuf<float4> planes;
  // 2-D, MFMA form
  DevBuf<int4> moffs;
  DevBuf<uint4> bmat;
  DevBuf<StencilGroup> mgroups;
  std::vector<StencilGroup> mgroups_h;  // host copies (MFMA instruction count)
  std::vector<int4> moffs_h;
  DevBuf<int> amb;
  DevBuf<uint4> planes_h;
  DevBuf<unsigned> sp_flags;
  DevBuf<float> mscale;
  // tile shapes: ev2.variant is the one-event (latency) shape; batches use
  // bvariant[path], the throughput shape, set by the autotuner when enabled
  int bvariant[2] = {0, 0};
  bool tuned[2] = {false, false};
  float tune_ms[2][3] = {{0, 0, 0}, {0, 0, 0}};
  // 3-D
  Event3D ev3{};
  DevBuf<RapSlot> slots;
  DevBuf<Samp3D> samp;
  DevBuf<int> eta_order;
  DevBuf<int2> srange;
  DevBuf<int4> win3;  // staged-row windows of the 3-D stream kernel
  std::vector<float> eta_work;  // kept cell-samples of each eta plane (work estimate)
  // (cell, sample) pairs the reference's masks keep, and (3-D) the pairs the
  // stream kernel evaluates (whole tiles per sample whose rows it reaches)
  double samples_kept = 0.0, samples_evaluated = 0.0;
  DevBuf<float> src;
  DevBuf<float> Tem;  // eta-major T
  // charge pass (BARYON = 1): normalisation of convertInitialChargeDensity;
  // 3-D: rapidity slots with the charge rapidity factors, source planes, J
  float charge_norm = 0.0f;
  Event3D ev3b{};
  DevBuf<RapSlot> slots_b;
  DevBuf<float> src_b;
  DevBuf<float> Jem;  // eta-major J (rows 0-3 of a 10-plane T)
  // device bytes held by the event-sized buffers (the plan cache's budget)
  size_t event_bytes() const {
    return planes.n * sizeof(float4) + planes_h.n * sizeof(uint4) + src.n * sizeof(float) + Tem.n * sizeof(float) +
           src_b.n * sizeof(float) + Jem.n * sizeof(float);
  }
};

// Plan cache limits: a batch of many distinct event shapes (every event of a
// mixed batch may have its own DX, DTAU and N_phi) keeps one plan each.
const size_t kPlanCacheMax = 1024;
const size_t kPlanCacheBytes = (size_t)8 << 30;

// src/FreeStream.cpp:365-366 with the reference's float/double mixing
// (SIGMA_B * SIGMA_B is a float product).
float charge_norm_factor(float SB) {
  const float sb2 = SB * SB;
  const float n = (float)(sqrt(2.0 * M_PI) * (double)SB * exp((double)sb2 / 2.0));
  return (float)(1.0 / (2.0 * M_PI * (double)n));
}

bool same_shape(const fsm_params &a, const fsm_params &b) { return memcmp(&a, &b, sizeof(a)) == 0; }

// Gauss-Hermite rule for the weight exp(-x^2) from the three-term recurrence
// of the orthonormal Hermite polynomials
//   p_0 = pi^(-1/4), p_{j+1} = x sqrt(2/(j+1)) p_j - sqrt(j/(j+1)) p_{j-1}:
// the nodes are the zeros of p_n (all inside |x| < sqrt(2n+1)), bracketed by
// sign changes on a grid finer than their spacing and refined by bisection;
// the weights are the Christoffel numbers 1 / sum_{j<n} p_j(x)^2.
void hermite_orthonormal(int n, double x, double *pn, double *sum_sq) {
  double pm = 0.0, p = 0.7511255444649425;  // pi^(-1/4)
  double ss = 0.0;
  for (int j = 0; j < n; ++j) {
    ss += p * p;
    const double pj = x * sqrt(2.0 / (j + 1)) * p - sqrt((double)j / (j + 1)) * pm;
    pm = p;
    p = pj;
  }
  *pn = p;
  *sum_sq = ss;
}

int gauss_hermite_rule(int n, double *x, double *w) {
  if (n < 1 || n > 200 || !x || !w) return FSM_ERR_ARG;
  const double L = sqrt(2.0 * n + 1.0) + 1.0;
  const double h = 0.25 / sqrt(2.0 * n + 1.0);  // zero spacing is at least ~pi / sqrt(2n+1)
  int found = 0;
  double a = -L, fa, ss;
  hermite_orthonormal(n, a, &fa, &ss);
  while (a < L && found < n) {
    const double b = std::min(a + h, L);
    double fb;
    hermite_orthonormal(n, b, &fb, &ss);
    if (fb == 0.0 || (fa < 0.0) != (fb < 0.0)) {
      double lo = a, hi = b, flo = fa;
      for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(1.0, fabs(lo)); ++it) {
        const double mid = 0.5 * (lo + hi);
        double fm;
        hermite_orthonormal(n, mid, &fm, &ss);
        if ((fm < 0.0) == (flo < 0.0)) {
          lo = mid;
          flo = fm;
        } else {
          hi = mid;
        }
      }
      x[found++] = 0.5 * (lo + hi);
      a = b;
      hermite_orthonormal(n, a, &fa, &ss);
      if (fb == 0.0) a += 1e-12;
      continue;
    }
    a = b;
    fa = fb;
  }
  if (found != n) return FSM_ERR_UNSUPPORTED;
  for (int k = 0; k < n; ++k) {
    // symmetric rule: the two halves agree to rounding; average them
    const double xs = 0.5 * (x[k] - x[n - 1 - k]);
    double pn;
    hermite_orthonormal(n, xs, &pn, &ss);
    x[k] = xs;
    w[k] = 1.0 / ss;
  }
  return FSM_OK;
}

void set_variant(Event2D &ev, int v) {
  ev.variant = v;
  ev.tiles_x = (ev.nx + fsm_stream2d_tile_w(v) - 1) / fsm_stream2d_tile_w(v);
  ev.tiles_y = (ev.ny + fsm_stream2d_tile_h(v) - 1) / fsm_stream2d_tile_h(v);
}

}  // namespace

// Workspace of fsm_run_device_noalloc: a fixed device array instead of
// allocations (one copy per GPU, alive until the process ends).  Host code
// uses only the address hipGetSymbolAddress returns for it, kept per device.
#define FSM_NOALLOC_SCRATCH_BYTES ((size_t)80 << 20)
#define FSM_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[FSM_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[FSM_NOALLOC_MAX_DEVICES];

struct fsm_ctx {
  int device = 0;
  int path = FSM_PATH_FP32;
  int rapidity_rule = FSM_RAPIDITY_TAN;  // 3-D events: fsm_set_rapidity_rule
  int tile_policy = FSM_TILES_AUTO;
  int autotune = 1;
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
  std::mutex tune_mu;  // one autotuning run at a time
  // host-API workspace
  DevBuf<float> d_in, d_nb, d_T, d_J;
  DevBuf<double> d_out;
  DevBuf<float> d_stage;  // planes of the stage entry points on host arrays
  // One-event calls share the workspace above and the plan's event-sized
  // buffers (spline planes, MFMA planes, 3-D source planes, eta-major T), so
  // they are ordered: a call's stream first waits for the previous call's
  // work (on whatever stream it was queued), and host threads take turns.
  std::recursive_mutex call_mu;  // fsm_run holds it around its nested fsm_run_device
  hipEvent_t last_call = nullptr;
  bool last_call_recorded = false;
  // device address of `scratch` on this context's device (null if the lookup failed)
  void *noalloc_scratch = nullptr;
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
  // 2-D charge pass: the reference builds the 2-D spline from rapidity copy 0
  // only (src/FreeStream.cpp:121-153), which for DIM_RAP > 1 differs from the
  // per-copy charge densities; only DIM_RAP = 1 is supported there
  if (p.BARYON && p.DIM_ETA == 1 && p.DIM_RAP != 1) return FSM_ERR_UNSUPPORTED;
  return FSM_OK;
}

// ---- 2-D plan ---------------------------------------------------------------
// Floor index of a shifted grid point as the reference computes it in float
// (src/FreeStream.cpp:163,187,198,204): x = (float)i*D + xmin, x_new = x -
// shift, floor((x_new - xmin)/D).  volatile keeps each rounding step.
int ref_floor(int i, float shift, float mn, float h) {
  volatile float x = (float)i * h;
  x = x + mn;
  volatile float xn = x - shift;
  volatile float d = xn - mn;
  volatile float q = d / h;
  return (int)floorf(q);
}

// True when the reference's floor index is i + o for every i in [0, n).
bool floor_is_offset(int n, float shift, float mn, float h, int o) {
  for (int i = 0; i < n; ++i)
    if (ref_floor(i, shift, mn, h) != i + o) return false;
  return true;
}

unsigned short half_bits(double v) {
  const _Float16 h = (_Float16)v;
  return __builtin_bit_cast(unsigned short, h);
}

// Tables of the MFMA form: distinct knot offsets of the samples selected by
// `use`, their 32x16 weight matrices in MFMA lane order,
// offset groups whose knot box (for a 32-wide tile) plus weights and offset
// records fit the LDS budget, and the remaining samples.  `st` is in phi
// order.  Offset record: (ox, oy, LDS delta in the group's box, 0).
struct MfmaTables {
  std::vector<int4> offs;
  std::vector<int4> recs;    // per group: its offsets' records, then its pair-MFMA records
  std::vector<uint4> bmat;   // per group, in the order of recs: 64 lanes per MFMA
  int npairs = 0;            // pair MFMAs of all groups (interior form)
  std::vector<StencilGroup> groups;
  int lds4 = 0;
  int rbx0 = 0, rbx1 = 0, rby0 = 0, rby1 = 0;
};

bool sample_offset_exact(const Stencil2D &S, int nx, int ny, float xmin, float ymin, float dx, float dy) {
  return floor_is_offset(nx, S.dxj, xmin, dx, S.ox) && floor_is_offset(ny, S.dyj, ymin, dy, S.oy);
}

void build_mfma_tables(const std::vector<Stencil2D> &st, const std::vector<char> &use, int budget4, MfmaTables &t) {
  std::vector<std::array<double, 96>> W;  // [corner*4 + comp][moment]
  for (size_t j = 0; j < st.size(); ++j) {
    if (!use[j]) continue;
    const Stencil2D &S = st[j];
    size_t o = 0;
    while (o < t.offs.size() && !(t.offs[o].x == S.ox && t.offs[o].y == S.oy)) ++o;
    if (o == t.offs.size()) {
      t.offs.push_back(make_int4(S.ox, S.oy, 0, 0));
      W.emplace_back();
      W.back().fill(0.0);
    }
    const double P[6] = {1.0, S.c, S.s, S.cc, S.cs, S.ss};
    for (int g = 0; g < 4; ++g) {
      const int gx = g & 1, gy = g >> 1;
      const double w[4] = {(double)S.hx[2 * gx] * S.hy[2 * gy], (double)S.hx[2 * gx] * S.hy[2 * gy + 1],
                           (double)S.hx[2 * gx + 1] * S.hy[2 * gy], (double)S.hx[2 * gx + 1] * S.hy[2 * gy + 1]};
      for (int c = 0; c < 4; ++c)
        for (int m = 0; m < 6; ++m) W[o][(g * 4 + c) * 6 + m] += w[c] * P[m];
    }
  }
  const size_t no = t.offs.size();
  // B operand of one MFMA from its 32x6 weights wt(g, c, m) (corner/lane
  // group g, knot component c, moment m): columns 0-5 hold W_h, columns 8-13
  // W_l, both for the hi and the lo knot halves.  Stored packed: lane l's
  // four distinct halves (W_h or W_l of its column for c = 0..3), 8 bytes, 64
  // lanes in 32 uint4; the kernel repeats them in the upper halves
  // (fsm_stream2d.hip, bmat_expand).
  auto bmat_lanes = [](auto wt, uint4 *out) {
    unsigned packed[128];
    for (int lane = 0; lane < 64; ++lane) {
      const int n = lane & 15, g = lane >> 4;
      unsigned short h[4] = {0, 0, 0, 0};
      for (int c = 0; c < 4; ++c) {
        if (n < 6) {
          h[c] = half_bits(wt(g, c, n));
        } else if (n >= 8 && n < 14) {
          const double w = wt(g, c, n - 8);
          const _Float16 hi = (_Float16)w;
          h[c] = half_bits(w - (double)hi);
        }
      }
      packed[2 * lane] = h[0] | ((unsigned)h[1] << 16);
      packed[2 * lane + 1] = h[2] | ((unsigned)h[3] << 16);
    }
    memcpy(out, packed, sizeof(packed));
  };
  // Interior form: the knot displacements d = o + corner of offsets [o0, o1)
  // (weights summed per d), covered row by row by dominoes (dx, dx+1) inside
  // the box [bx0, bx1] (a displacement at bx1 with no partner takes the
  // domino (bx1-1, bx1)); each displacement's weights go to exactly one half.
  struct Domino {
    int dx, dy;
    std::array<double, 24> v[2];  // halves gx = 0, 1: [c * 6 + m]
  };
  auto cover = [&](size_t o0, size_t o1, int bx1, std::vector<Domino> *out) {
    std::map<std::pair<int, int>, std::array<double, 24>> V;  // (dy, dx) -> weights
    for (size_t o = o0; o < o1; ++o)
      for (int g = 0; g < 4; ++g) {
        auto &v = V[{t.offs[o].y + (g >> 1), t.offs[o].x + (g & 1)}];
        for (int k = 0; k < 24; ++k) v[k] += W[o][g * 24 + k];
      }
    int count = 0;
    for (auto it = V.begin(); it != V.end();) {
      Domino D;
      D.dy = it->first.first;
      const int d = it->first.second;
      D.v[0].fill(0.0);
      D.v[1].fill(0.0);
      if (d + 1 <= bx1) {
        D.dx = d;
        D.v[0] = it->second;
        ++it;
        if (it != V.end() && it->first.first == D.dy && it->first.second == d + 1) {
          D.v[1] = it->second;
          ++it;
        }
      } else {
        D.dx = d - 1;
        D.v[1] = it->second;
        ++it;
      }
      ++count;
      if (out) out->push_back(D);
    }
    return count;
  };
  std::vector<std::vector<uint4>> offb(no, std::vector<uint4>(FSM_BMAT_U4));
  for (size_t o = 0; o < no; ++o)
    bmat_lanes([&](int g, int c, int m) { return W[o][(g * 4 + c) * 6 + m]; }, offb[o].data());
  // greedy groups along the phi order of first appearance; a group's LDS
  // image is its knot box and one record (16 B) per MFMA of either form (the
  // packed B operands are read from global memory)
  const int TX = fsm_stream2d_tile_w(0), TY = fsm_stream2d_tile_h(0);
  t.lds4 = 0;
  t.rbx0 = t.rby0 = 1 << 30;
  t.rbx1 = t.rby1 = -(1 << 30);
  t.bmat.clear();
  std::vector<int4> recs;
  size_t o = 0;
  while (o < no) {
    StencilGroup G{};
    G.j0 = (int)o;
    G.bx0 = t.offs[o].x;
    G.bx1 = t.offs[o].x + 1;
    G.by0 = t.offs[o].y;
    G.by1 = t.offs[o].y + 1;
    size_t e = o + 1;
    for (; e < no; ++e) {
      const int bx0 = std::min(G.bx0, t.offs[e].x), bx1 = std::max(G.bx1, t.offs[e].x + 1);
      const int by0 = std::min(G.by0, t.offs[e].y), by1 = std::max(G.by1, t.offs[e].y + 1);
      const int np = (cover(o, e + 1, bx1, nullptr) + 1) / 2;
      if ((TX + bx1 - bx0) * ((TY + by1 - by0 + 15) & ~15) + ((int)(e + 1 - o) + np) > budget4)
        break;
      G.bx0 = bx0;
      G.bx1 = bx1;
      G.by0 = by0;
      G.by1 = by1;
    }
    G.j1 = (int)e;
    std::vector<Domino> dom;
    cover(o, e, G.bx1, &dom);
    const int np = ((int)dom.size() + 1) / 2;
    // knot box columns padded to 16 records (TY rows + halo, rounded up)
    const int LYP = (TY + G.by1 - G.by0 + 15) & ~15;
    t.lds4 = std::max(t.lds4, (TX + G.bx1 - G.bx0) * LYP + (G.j1 - G.j0 + np));
    // the group's block: its offsets' records and B operands, then its pairs'
    G.pad0 = (int)recs.size();
    G.pad1 = np;
    for (size_t k = o; k < e; ++k) {
      // LDS delta of each offset inside the group's knot box
      t.offs[k].z = (t.offs[k].x - G.bx0) * LYP + (t.offs[k].y - G.by0);
      recs.push_back(t.offs[k]);
      t.bmat.insert(t.bmat.end(), offb[k].begin(), offb[k].end());
    }
    for (int q = 0; q < np; ++q) {
      // lane groups gy = 0, 1 take dominoes 2q, 2q+1 (an odd count leaves the
      // second half empty, reading the first domino's knots with zero weights)
      const Domino &A = dom[2 * q];
      const bool two = 2 * q + 1 < (int)dom.size();
      const Domino &Bd = two ? dom[2 * q + 1] : A;
      const int da = (A.dx - G.bx0) * LYP + (A.dy - G.by0), db = (Bd.dx - G.bx0) * LYP + (Bd.dy - G.by0);
      recs.push_back(make_int4(da, db, 0, 0));
      uint4 lanes[FSM_BMAT_U4];
      bmat_lanes(
          [&](int g, int c, int m) {
            const int gx = g & 1, gy = g >> 1;
            if (gy && !two) return 0.0;
            return (gy ? Bd : A).v[gx][c * 6 + m];
          },
          lanes);
      t.bmat.insert(t.bmat.end(), lanes, lanes + FSM_BMAT_U4);
    }
    t.rbx0 = std::min(t.rbx0, G.bx0);
    t.rbx1 = std::max(t.rbx1, G.bx1);
    t.rby0 = std::min(t.rby0, G.by0);
    t.rby1 = std::max(t.rby1, G.by1);
    t.groups.push_back(G);
    t.npairs += np;
    o = e;
  }
  t.recs = recs;
  if (t.bmat.empty()) t.bmat.assign(FSM_BMAT_U4, make_uint4(0, 0, 0, 0));
  if (t.recs.empty()) t.recs.push_back(make_int4(0, 0, 0, 0));
  if (no == 0) t.rbx0 = t.rbx1 = t.rby0 = t.rby1 = 0;
}

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
  {
    // cells whose sample j the reference keeps: floor index ix + ox in [1, nx-2]
    // (a sample whose shift is within rounding of a whole cell may differ by
    // one row of cells; that does not matter for a work count)
    auto kept_run = [](int n, int o) { return (double)std::max(0, std::min(n - 2 - o, n - 1) - std::max(1 - o, 0) + 1); };
    pl.samples_kept = 0.0;
    for (int j = 0; j < nphi; ++j) pl.samples_kept += kept_run(nx, st[j].ox) * kept_run(ny, st[j].oy);
    pl.samples_kept *= p.DIM_RAP;  // identical rapidity copies in 2-D
    pl.samples_evaluated = (double)nx * ny * nphi;
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
  const float xmin = (float)((-1.0) * ((double)(float)(nx - 1) / 2.0) * (double)p.DX);
  const float ymin = (float)((-1.0) * ((double)(float)(ny - 1) / 2.0) * (double)p.DY);
  // every sample enters W; the rounding-dependent ones (amb below) also get
  // an edge-mask correction on the VALU
  std::vector<char> all(nphi, 1);
  MfmaTables mt;
  build_mfma_tables(st, all, budget4, mt);
  // Within a group the order of the samples is free (it only changes the
  // summation order): sort by knot offset so equal offsets are adjacent.
  for (const StencilGroup &sg : groups)
    std::stable_sort(st.begin() + sg.j0, st.begin() + sg.j1, [](const Stencil2D &a, const Stencil2D &b) {
      return a.ox != b.ox ? a.ox < b.ox : a.oy < b.oy;
    });
  std::vector<int> amb;
  for (int j = 0; j < nphi; ++j)
    if (!sample_offset_exact(st[j], nx, ny, xmin, ymin, p.DX, p.DY)) amb.push_back(j);
  FSM_TRY(pl.gx.upload(gxw));
  FSM_TRY(pl.gy.upload(gyw));
  // interior rows (columns) whose band weights are bitwise those of the middle row
  auto canon = [](const std::vector<float> &g, int n, int *lo, int *hi) {
    const int W = FSM_SPLINE_W, c = n / 2;
    auto same = [&](int i) { return memcmp(&g[(size_t)i * W], &g[(size_t)c * W], W * sizeof(float)) == 0; };
    *lo = *hi = c;
    while (*lo > 0 && same(*lo - 1)) --*lo;
    while (*hi < n - 1 && same(*hi + 1)) ++*hi;
  };
  int cxlo, cxhi, cylo, cyhi;
  canon(gxw, nx, &cxlo, &cxhi);
  canon(gyw, ny, &cylo, &cyhi);
  FSM_TRY(pl.st.upload(st));
  FSM_TRY(pl.groups.upload(groups));
  FSM_TRY(pl.planes.ensure((size_t)nx * ny));
  FSM_TRY(pl.moffs.upload(mt.recs));
  FSM_TRY(pl.bmat.upload(mt.bmat));
  FSM_TRY(pl.mgroups.upload(mt.groups));
  pl.mgroups_h = mt.groups;
  pl.moffs_h = mt.offs;
  FSM_TRY(pl.amb.upload(amb));
  const int stx = (nx + FSMSP_TX - 1) / FSMSP_TX, sty = (ny + FSMSP_TY - 1) / FSMSP_TY;
  FSM_TRY(pl.planes_h.ensure((size_t)nx * ny));
  FSM_TRY(pl.sp_flags.ensure((size_t)stx * sty));
  FSM_TRY(pl.mscale.ensure(FSM_KS_PARTS + 1));
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
  // two workgroups per CU; batches start from the widest shape (bvariant).
  int sv = 2;
  for (int v = 0; v < 3; ++v) {
    int tw = fsm_stream2d_tile_w(v), th = fsm_stream2d_tile_h(v);
    long tiles = (long)((nx + tw - 1) / tw) * ((ny + th - 1) / th);
    if (tiles >= 2L * ctx->num_cus) {
      sv = v;
      break;
    }
  }
  set_variant(ev, sv);
  ev.xmin = xmin;
  ev.ymin = ymin;
  ev.dx = p.DX;
  ev.dy = p.DY;
  ev.planes = pl.planes.p;
  ev.gx = pl.gx.p;
  ev.gy = pl.gy.p;
  ev.cxlo = cxlo;
  ev.cxhi = cxhi;
  ev.cylo = cylo;
  ev.cyhi = cyhi;
  ev.st = pl.st.p;
  ev.groups = pl.groups.p;
  ev.t_plane = p.DIM;
  ev.t_row = ny;
  ev.mfma = 0;  // set per call from the context's path
  ev.n_moff = (int)mt.offs.size();
  ev.n_mgroups = (int)mt.groups.size();
  ev.n_amb = (int)amb.size();
  ev.mfma_lds4 = mt.lds4;
  ev.rbx0 = mt.rbx0;
  ev.rbx1 = mt.rbx1;
  ev.rby0 = mt.rby0;
  ev.rby1 = mt.rby1;
  ev.sp_tiles_y = sty;
  ev.planes_h = nullptr;
  ev.sp_flags = pl.sp_flags.p;
  ev.mscale = pl.mscale.p;
  pl.charge_norm = charge_norm_factor(p.SIGMA_B);
  ev.charge = 0;
  ev.raw_e = p.IC_ENERGY != 5;
  ev.in_norm = pl.charge_norm;
  ev.out_j = 0;
  ev.jscale = p.DRAP;
  ev.moffs = pl.moffs.p;
  ev.bmat = pl.bmat.p;
  ev.mgroups = pl.mgroups.p;
  ev.amb = pl.amb.p;
  return FSM_OK;
}

// ---- 3-D plan ---------------------------------------------------------------
// Floor of the reference's float grid index of a shifted sample
// (src/FreeStream.cpp:163,193,198,204): x = (float)i*h + mn, x_new = x -
// shift, floor((x_new - mn)/h); no contraction, as in the reference build.
float ref_floor_index(int i, float shift, float mn, float h) {
#pragma clang fp contract(off)
  const float x = (float)i * h + mn;
  const float xn = x - shift;
  return floorf((xn - mn) / h);
}

// Cells i in [0, n) whose sample the reference keeps along one axis
// (floor index in [1, n-2], src/FreeStream.cpp:213).  The floor index is
// monotone in i, so they form one interval [*lo, *hi] (empty: *lo > *hi); g0
// is the cell-independent floor offset, within one of the exact floor.
void kept_interval(int n, int g0, float shift, float mn, float h, int *lo, int *hi) {
  *lo = 1;
  *hi = 0;
  int i = std::max(0, 1 - g0 - 2);
  if (i > n - 1) return;
  while (i < n && ref_floor_index(i, shift, mn, h) < 1.0f) ++i;
  int k = std::min(n - 1, n - 2 - g0 + 2);
  if (k < 0) return;
  while (k >= 0 && !(ref_floor_index(k, shift, mn, h) < (float)(n - 1))) --k;
  if (i > k) return;
  *lo = i;
  *hi = k;
}

// One azimuth sample (cos c, sin s) of a slot with shift magnitude D:
// offsets, bilinear weights and the kept intervals.  Returns 0 with
// *kept = false when the sample keeps no cell, -1 if a kept cell would read
// outside [0, N] (cannot happen for grids of sane size; the planner refuses
// the shape rather than read out of bounds).
int make_samp3d(float c, float s, float D, const fsm_params &p, Samp3D *P, bool *kept) {
#pragma clang fp contract(off)
  const int nx = p.DIM_X, ny = p.DIM_Y;
  const float xmin = (float)((-1.0) * ((double)(float)(nx - 1) / 2.0) * (double)p.DX);
  const float ymin = (float)((-1.0) * ((double)(float)(ny - 1) / 2.0) * (double)p.DY);
  // src/FreeStream.cpp:193-194: x_new = x - cos(phip) * D
  const float dxj = c * D, dyj = s * D;
  const float sx = -dxj * (1.0f / p.DX), sy = -dyj * (1.0f / p.DY);
  const float fx = floorf(sx), fy = floorf(sy);
  memset(P, 0, sizeof(*P));
  *kept = false;
  // a sample shifted by more than the grid keeps nothing
  if (!(fabsf(sx) < (float)(nx + 2)) || !(fabsf(sy) < (float)(ny + 2))) return 0;
  P->ox = (int)fx;
  P->oy = (int)fy;
  const float tx = sx - fx, ty = sy - fy;
  int ylo = 1, yhi = 0;
  kept_interval(nx, P->ox, dxj, xmin, p.DX, &P->xlo, &P->xhi);
  kept_interval(ny, P->oy, dyj, ymin, p.DY, &ylo, &yhi);
  if (P->xlo > P->xhi || ylo > yhi) return 0;
  if (P->xlo + P->ox < 0 || P->xhi + P->ox + 1 > nx || ylo + P->oy < 0 || yhi + P->oy + 1 > ny) return -1;
  *kept = true;
  P->mlo = ((unsigned)P->xlo << 16) | (unsigned)ylo;
  P->mext = ((unsigned)(P->xhi - P->xlo) << 16) | (unsigned)(yhi - ylo);
  P->d0 = 4 * (P->ox * ny + P->oy);
  P->w00 = (1.0f - tx) * (1.0f - ty);
  P->w01 = (1.0f - tx) * ty;
  P->w10 = tx * (1.0f - ty);
  P->w11 = tx * ty;
  P->c = c;
  P->s = s;
  P->cc = c * c;
  P->cs = c * s;
  return 0;
}

// Kept cell-samples of one 3-D sample (the product of its kept intervals).
double samp3d_kept(const Samp3D &P) {
  return (double)((P.mext >> 16) + 1) * (double)((P.mext & 0xFFFFu) + 1);
}

// Staged-row windows of one (slot, tile) pair for the 3-D stream kernel.
// The tile is cells [c0, c0 + FSM3D_TILE) of the ix * NY + iy order, rows
// [rlo, rhi].  Walking the slot's samples in their (ox, oy) order, a window
// starts at the first sample that keeps a cell in the tile's rows and grows
// over the following samples while the source rows their kept cells read fit
// W rows; the sample that does not fit starts the next window.  Samples
// keeping no cell of the tile's rows are left out (in the (ox, oy) order they
// lie before or after the tile's samples).  Each window is (first sample,
// end sample, first source row, rows).
void plan_windows3d(const Samp3D *sp, int ns, int rlo, int rhi, int W, std::vector<int4> &out) {
  int j = 0;
  while (j < ns) {
    const Samp3D &P = sp[j];
    const int a = std::max(rlo, P.xlo), b = std::min(rhi, P.xhi);
    if (a > b) {
      ++j;
      continue;
    }
    int lo = a + P.ox, hi = b + P.ox + 1, end = j + 1, j2 = j + 1;
    for (; j2 < ns; ++j2) {
      const Samp3D &Q = sp[j2];
      const int a2 = std::max(rlo, Q.xlo), b2 = std::min(rhi, Q.xhi);
      if (a2 > b2) continue;
      const int l2 = std::min(lo, a2 + Q.ox), h2 = std::max(hi, b2 + Q.ox + 1);
      if (h2 - l2 + 1 > W) break;
      lo = l2;
      hi = h2;
      end = j2 + 1;
    }
    out.push_back(make_int4(j, end, lo, hi - lo + 1));
    j = j2;
  }
}

int build_plan3d(fsm_ctx *ctx, const fsm_params &p, Plan &pl) {
#pragma clang fp contract(off)
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
  const bool gh = ctx->rapidity_rule == FSM_RAPIDITY_GAUSS_HERMITE;
  std::vector<double> ghx(nrap), ghw(nrap);
  if (gh) {
    const int st = gauss_hermite_rule(nrap, ghx.data(), ghw.data());
    if (st) return st;
  }
  auto wk = [&](int k) { return (float)(-.9975 + (double)(float)k * (1.995 / (double)(float)(nrap - 1))); };
  // rapidity node k of a cell at eta: the reference's tan grid
  // (src/FreeStream.cpp:174-175), or eta + sqrt(2) SIGMA x_k
  auto rap_of = [&](float eta, int k) {
    if (gh) return (float)((double)eta + M_SQRT2 * (double)SIGMA * ghx[k]);
    return (float)((double)eta + tan((M_PI / 2.0) * (double)wk(k)));
  };
  // rapidity factor of convertInitialDensity (src/FreeStream.cpp:342-345);
  // under the Gauss-Hermite rule the Gaussian is in the weights
  auto rfac = [&](int ie, int k) {
    float eta = (float)ie * DETA + etamin;
    float rap = rap_of(eta, k);
    float d = eta - rap;
    if (gh) return (float)((double)(coshf(d) * coshf(d)));
    return (float)((double)(coshf(d) * coshf(d)) *
                   exp((-1.0) * (double)d * (double)d / (2.0 * (double)SIGMA * (double)SIGMA)));
  };
  std::vector<RapSlot> slots((size_t)neta * nrap);
  std::vector<Samp3D> samp;
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
        S.jac = gh ? (float)(M_SQRT2 * (double)SIGMA * ghw[k])
                   : (float)((M_PI / 2.0) / cos((M_PI / 2.0) * (double)w) / cos((M_PI / 2.0) * (double)w) *
                             (1.995 / (double)(float)(nrap - 1)));
        const float ch = coshf(rap - eta), th = tanhf(rap - eta);
        S.inv_ch = 1.0f / ch;
        S.inv_ch2 = 1.0f / (ch * ch);
        S.th_t = (float)((1.0 / (double)TAU) * (double)th);
        S.th_tch = (float)((1.0 / (double)TAU) * (double)th / (double)ch);
        S.th2_t2 = (float)((1.0 / (double)(TAU * TAU)) * (double)th * (double)th);
        S.s0 = (int)samp.size();
        for (int j = 0; j < nphi; ++j) {
          Samp3D P;
          bool kept = false;
          if (make_samp3d(cs[2 * j], cs[2 * j + 1], D, p, &P, &kept) < 0) return FSM_ERR_UNSUPPORTED;
          if (kept) samp.push_back(P);
        }
        S.ns = (int)samp.size() - S.s0;
        // by source row offset: a tile's samples form one index range and
        // consecutive samples read neighbouring rows
        std::stable_sort(samp.begin() + S.s0, samp.end(), [](const Samp3D &u, const Samp3D &v) {
          return u.ox < v.ox || (u.ox == v.ox && u.oy < v.oy);
        });
      }
      slots[(size_t)ie * nrap + k] = S;
    }
  }
  FSM_TRY(pl.slots.upload(slots));
  // work per eta plane: kept cell-samples of its slots; the one-event launch
  // starts the heaviest planes first
  pl.eta_work.assign(neta, 0.0f);
  pl.samples_kept = 0.0;
  for (int ie = 0; ie < neta; ++ie)
    for (int k = 0; k < nrap; ++k) {
      const RapSlot &S = slots[(size_t)ie * nrap + k];
      for (int j = S.s0; j < S.s0 + S.ns; ++j) {
        const double kept = samp3d_kept(samp[j]);
        pl.eta_work[ie] += (float)kept;
        pl.samples_kept += kept;
      }
    }
  {
    std::vector<int> order(neta);
    for (int ie = 0; ie < neta; ++ie) order[ie] = ie;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return pl.eta_work[x] > pl.eta_work[y]; });
    FSM_TRY(pl.eta_order.upload(order));
  }
  // staged-row window of the stream kernel: at least the rows one sample of a
  // tile reads (the tile's rows + 1), else as many as fit FSM3D_WIN_FLOATS
  int win_rows;
  {
    const int need = (FSM3D_TILE - 1) / ny + 3;
    win_rows = std::max(need, FSM3D_WIN_FLOATS / ny);
    win_rows = std::min(win_rows, nx + 2);
    win_rows = std::max(win_rows, need);
  }
  {
    // per (slot, tile): the staged-row windows of the slot's samples that keep
    // a cell in the tile's rows (tiles are FSM3D_TILE consecutive cells)
    const int tiles = (nx * ny + FSM3D_TILE - 1) / FSM3D_TILE;
    std::vector<int2> rg((size_t)neta * nrap * tiles, make_int2(0, 0));
    std::vector<int4> wins;
    pl.samples_evaluated = 0.0;
    for (size_t si = 0; si < slots.size(); ++si) {
      const RapSlot &S = slots[si];
      if (!S.valid || S.ns == 0) continue;
      for (int t = 0; t < tiles; ++t) {
        const int c0 = t * FSM3D_TILE;
        const int rlo = c0 / ny, rhi = std::min(c0 + FSM3D_TILE - 1, nx * ny - 1) / ny;
        const size_t w0 = wins.size();
        plan_windows3d(samp.data() + S.s0, S.ns, rlo, rhi, win_rows, wins);
        for (size_t w = w0; w < wins.size(); ++w)
          pl.samples_evaluated += (double)FSM3D_TILE * (wins[w].y - wins[w].x);  // whole tiles per sample
        if (wins.size() > w0) rg[si * tiles + t] = make_int2((int)w0, (int)wins.size());
      }
    }
    if (wins.empty()) wins.push_back(make_int4(0, 0, 0, 0));  // keep a valid pointer
    FSM_TRY(pl.srange.upload(rg));
    FSM_TRY(pl.win3.upload(wins));
  }
  if (samp.empty()) samp.push_back(Samp3D{});  // keep a valid pointer
  FSM_TRY(pl.samp.upload(samp));
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
  ev.win_rows = win_rows;
  ev.lds_floats = (win_rows + 2) * ny;  // + the 2 x 2 zero patch read by dropped cells
  if ((size_t)ev.lds_floats * sizeof(float) > (size_t)ctx->max_lds) return FSM_ERR_UNSUPPORTED;
  ev.d_phip = (float)((2.0 * M_PI) / (double)(float)nphi);
  ev.inv_dx = inv_dx;
  ev.inv_dy = inv_dy;
  ev.dx = p.DX;
  ev.dy = p.DY;
  ev.xmin = (float)((-1.0) * ((double)(float)(nx - 1) / 2.0) * (double)p.DX);
  ev.ymin = (float)((-1.0) * ((double)(float)(ny - 1) / 2.0) * (double)p.DY);
  ev.src = pl.src.p;
  ev.slots = pl.slots.p;
  ev.samp = pl.samp.p;
  ev.srange = pl.srange.p;
  ev.win = pl.win3.p;
  ev.eta_order = pl.eta_order.p;
  ev.T = pl.Tem.p;
  ev.t_plane = p.DIM;
  {
    double n = sqrt(M_PI / 2.0) * (double)SIGMA * (1.0 + exp(2.0 * (double)SIGMA * (double)SIGMA));
    float nf = (float)n;
    ev.norm_factor = (float)(1.0 / (2.0 * M_PI * (double)nf));
  }
  ev.charge = 0;
  ev.raw_e = p.IC_ENERGY != 5;
  if (p.BARYON) {
    // charge pass: same geometry; the density carries the charge rapidity
    // factor on the uniform DRAP grid (src/FreeStream.cpp:368-385) and
    // calculateBaryonCurrent weights every slot by DRAP (src/LandauMatch.cpp:140)
    const float SB = p.SIGMA_B, DRAP = p.DRAP, sb2 = SB * SB;
    const float rapmin = (float)((-1.0) * ((double)(float)(nrap - 1) / 2.0) * (double)DRAP);
    auto crf = [&](int ie, int k) {
      const float eta = (float)ie * DETA + etamin;
      const float rap = (float)k * DRAP + rapmin;
      const float d = eta - rap;
      return (float)((double)coshf(d) * exp((-1.0) * (double)d * (double)d / (double)sb2));
    };
    std::vector<RapSlot> sb(slots);
    for (int ie = 0; ie < neta; ++ie)
      for (int k = 0; k < nrap; ++k) {
        RapSlot &S = sb[(size_t)ie * nrap + k];
        if (!S.valid) continue;
        S.rf0 = crf(S.eta0, k);
        S.rf1 = crf(S.eta0 + 1, k);
        S.jac = DRAP;
      }
    FSM_TRY(pl.slots_b.upload(sb));
    FSM_TRY(pl.src_b.ensure((size_t)p.DIM));
    FSM_TRY(pl.Jem.ensure((size_t)10 * p.DIM));
    pl.charge_norm = charge_norm_factor(SB);
    pl.ev3b = ev;
    pl.ev3b.charge = 1;
    pl.ev3b.norm_factor = pl.charge_norm;
    pl.ev3b.src = pl.src_b.p;
    pl.ev3b.slots = pl.slots_b.p;
    pl.ev3b.T = pl.Jem.p;
  }
  return FSM_OK;
}

// Plans are cached per parameter set (most recently used first).  A batch
// keeps shared references to the plans of its events, so eviction from the
// cache never frees tables a recorded batch still uses.
// A cached plan serves p on ctx: same parameters, and for 3-D events the
// context's rapidity rule.
bool plan_matches(const Plan &pl, const fsm_params &p, const fsm_ctx *ctx) {
  return same_shape(pl.key, p) && (!pl.is3d || pl.rapidity_rule == ctx->rapidity_rule);
}

int get_plan_shared(fsm_ctx *ctx, const fsm_params &p, std::shared_ptr<Plan> *out) {
  std::lock_guard<std::mutex> lock(ctx->plan_mu);
  for (size_t i = 0; i < ctx->plans.size(); ++i) {
    if (plan_matches(*ctx->plans[i], p, ctx)) {
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
  pl->rapidity_rule = pl->is3d ? ctx->rapidity_rule : FSM_RAPIDITY_TAN;
  st = pl->is3d ? build_plan3d(ctx, p, *pl) : build_plan2d(ctx, p, *pl);
  if (st) return st;
  pl->valid = true;
  ctx->plans.insert(ctx->plans.begin(), pl);
  size_t bytes = 0;
  for (auto &q : ctx->plans) bytes += q->event_bytes();
  while (ctx->plans.size() > 1 && (ctx->plans.size() > kPlanCacheMax || bytes > kPlanCacheBytes)) {
    bytes -= ctx->plans.back()->event_bytes();
    ctx->plans.pop_back();
  }
  *out = pl;
  return FSM_OK;
}

// Batched device time per unit of work on MI355X, in picoseconds, from the
// measured stage times of the 261^2 lumpy event (64-event batches: stream
// 14.4 / 8.5 us for 13.6 M samples, spline 1.5 us and Landau 3.0 us for 68 k
// cells), of the 131^2 x 101 event (stream 0.47 ms for 6.8e8 evaluated
// samples) and of a batch of four 131^2 x 41..101 events (Landau 0.27 ms for
// 4.87 M cells): 2-D stream per (cell, phi) sample on each path, 3-D stream
// per evaluated sample, spline or source planes and Landau per cell.
const double kCostStream2dPs[2] = {1.06, 0.62};  // FP32, mixed
const double kCostStream3dPs = 0.70;
const double kCostPlanes2dPs = 22.0, kCostPlanes3dPs = 17.0;
const double kCostLandau2dPs = 36.0, kCostLandau3dPs = 55.0;

// Estimated device time of one event inside a batch, microseconds.
double event_cost_us(const Plan &pl, const fsm_params &p, int path, bool baryon) {
  const int passes = baryon ? 2 : 1;  // energy and charge streaming
  const double D = (double)p.DIM;
  double ps;
  if (pl.is3d)
    ps = passes * (pl.samples_evaluated * kCostStream3dPs + D * kCostPlanes3dPs) + D * kCostLandau3dPs;
  else
    ps = passes * (pl.samples_evaluated * kCostStream2dPs[path == FSM_PATH_MIXED ? 1 : 0] + D * kCostPlanes2dPs) +
         D * kCostLandau2dPs;
  return ps * 1e-6;
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

EventLandau landau_event(const fsm_params &p, const float *T, int eta_major);  // below

// The 2-D stream kernels' fused Landau stage: the tile epilogue solves the
// 3x3 frame on the moments in registers and writes the 17 output planes of
// d_out (events without a baryon current); EventLandau's constants.
void fuse_landau2d(Event2D &ev, const fsm_params &p, double *d_out) {
  const EventLandau L = landau_event(p, nullptr, 0);
  ev.fuse_landau = 1;
  ev.l_eos = L.eos_type;
  ev.l_viscous = L.viscous;
  ev.l_fill = L.fill_outputs;
  ev.l_inv_tau2 = L.inv_tau2;
  ev.l_out = d_out;
  ev.l_out_plane = L.out_plane;
}

// Stage A into T with layout chosen by the plan (2-D: reference layout into
// T_out; 3-D: eta-major into the plan's buffer).  fuse_out (2-D only): run
// the Landau stage in the stream kernel's epilogue into these outputs
// instead of writing T.
int stage_a(fsm_ctx *ctx, Plan &pl, const fsm_params &p, const float *d_e, float *T_ref, hipStream_t s,
            const float **T_used, int *eta_major, double *fuse_out = nullptr) {
  if (!pl.is3d) {
    Event2D ev = pl.ev2;
    ev.e_gev = d_e;
    ev.T = T_ref;
    if (fuse_out) fuse_landau2d(ev, p, fuse_out);
    if (ctx->tile_policy == FSM_TILES_BATCH) set_variant(ev, pl.bvariant[ctx->path]);
    if (ctx->path == FSM_PATH_MIXED) {
      ev.mfma = 1;
      ev.planes_h = pl.planes_h.p;
    }
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

// Charge pass of the baryon extension: n_B through convertInitialChargeDensity
// and freeStream into J^mu (calculateBaryonCurrent), in the plan's buffers
// after the energy pass on the same stream.  2-D: J goes to J_ref (4 planes,
// reference layout); 3-D: to the plan's eta-major buffer.
int stream_charge(fsm_ctx *ctx, Plan &pl, const fsm_params &p, const float *d_nb, float *J_ref, hipStream_t s,
                  const float **J_used) {
  if (!pl.is3d) {
    Event2D ev = pl.ev2;
    ev.e_gev = d_nb;
    ev.T = J_ref;
    ev.charge = 1;
    ev.out_j = 1;
    ev.out_scale = 1.0f;  // DIM_RAP = 1
    if (ctx->tile_policy == FSM_TILES_BATCH) set_variant(ev, pl.bvariant[ctx->path]);
    if (ctx->path == FSM_PATH_MIXED) {
      ev.mfma = 1;
      ev.planes_h = pl.planes_h.p;
    }
    FSM_TRY(fsm_launch_spline2d(ev, s));
    FSM_TRY(fsm_launch_stream2d(ev, s));
    *J_used = J_ref;
  } else {
    Event3D ev = pl.ev3b;
    ev.e_gev = d_nb;
    FSM_TRY(fsm_launch_src3d(ev, s));
    FSM_TRY(fsm_launch_stream3d(ev, s));
    *J_used = pl.Jem.p;
  }
  (void)p;
  return FSM_OK;
}

EventLandau landau_event(const fsm_params &p, const float *T, int eta_major) {
  EventLandau L;
  memset(&L, 0, sizeof(L));
  L.dim = p.DIM;
  L.nxy = p.DIM_X * p.DIM_Y;
  L.neta = p.DIM_ETA;
  L.t_eta_major = eta_major;
  L.eos_type = p.EOS_TYPE;
  L.viscous = p.VISCOUS_MATCHING;
  L.fill_outputs = (p.OUTPUTFORMAT == 2 || p.OUTPUTFORMAT == 3) ? 1 : 0;
  L.tau = p.TAU;
  {
    const float t2 = p.TAU * p.TAU;  // a float product, as the reference forms it
    L.inv_tau2 = 1.0 / (double)t2;
  }
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

int autotune2d(fsm_ctx *ctx, Plan &pl, const fsm_params &p);  // batch section

// Orders a one-event call after the previous one on the same context (see
// fsm_ctx::call_mu).  A stream under graph capture is left alone: capture
// cannot wait on work recorded outside it, and the caller's graph orders
// its own nodes.
struct CallOrder {
  fsm_ctx *ctx;
  hipStream_t s;
  bool capturing = false;
  std::lock_guard<std::recursive_mutex> lock;
  CallOrder(fsm_ctx *c, hipStream_t st) : ctx(c), s(st), lock(c->call_mu) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (s && hipStreamIsCapturing(s, &cs) == hipSuccess) capturing = cs != hipStreamCaptureStatusNone;
  }
  hipError_t begin() {
    if (capturing || !ctx->last_call_recorded) return hipSuccess;
    return hipStreamWaitEvent(s, ctx->last_call, 0);
  }
  ~CallOrder() {
    if (capturing || !ctx->last_call) return;
    if (hipEventRecord(ctx->last_call, s) == hipSuccess) ctx->last_call_recorded = true;
  }
};

int run_device_impl(fsm_ctx *ctx, const fsm_params *p, const float *d_e, double *d_out, float *T_scratch,
                    hipStream_t s, const float *d_nb = nullptr, float *J_scratch = nullptr) {
  // held for the call: the server thread may evict it from the cache meanwhile
  std::shared_ptr<Plan> plan;
  int st = get_plan_shared(ctx, *p, &plan);
  if (st) return st;
  Plan *pl = plan.get();
  const float *T = nullptr;
  int em = 0;
  // 2-D events without a baryon current: Landau fused into the stream kernel
  const bool fuse = !pl->is3d && !d_nb;
  st = stage_a(ctx, *pl, *p, d_e, T_scratch, s, &T, &em, fuse ? d_out : nullptr);
  if (st) return st;
  if (fuse) {
    tmark(ctx, 3, s);
    return FSM_OK;
  }
  const float *J = nullptr;
  if (d_nb) {
    st = stream_charge(ctx, *pl, *p, d_nb, J_scratch, s, &J);
    if (st) return st;
  }
  EventLandau L = landau_event(*p, T, em);
  L.eta_zero = p->DIM_ETA == 1;  // the 2-D stream kernel writes the eta components as zeros
  L.out = d_out;
  if (J) {
    L.J = J;
    L.out_b = d_out + (size_t)FSM_NUM_OUTPUTS * p->DIM;
  }
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
    case FSM_ERR_TOO_LARGE: return "event exceeds the fixed workspace of fsm_run_device_noalloc";
    case FSM_ERR_NO_PLAN: return "event shape not prepared: no tables for it in this context yet";
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
  FSM_TRY(hipEventCreateWithFlags(&ctx->last_call, hipEventDisableTiming));
  if (device < FSM_NOALLOC_MAX_DEVICES) {
    // once per device, with that device current: the address of its copy of
    // the allocation-free entry point's workspace
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
      void *ptr = nullptr;
      if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_ptr[device] = ptr;
      else (void)hipGetLastError();  // fsm_run_device_noalloc then reports FSM_ERR_DEVICE
    }
    ctx->noalloc_scratch = g_scratch_ptr[device];
  }
  *out = ctx.release();
  return FSM_OK;
}

void fsm_destroy(fsm_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  for (auto &set : ctx->pool)
    for (auto &e : set) (void)hipEventDestroy(e);
  if (ctx->last_call) (void)hipEventDestroy(ctx->last_call);
  delete ctx;
}

int fsm_set_rapidity_rule(fsm_ctx *ctx, int rule) {
  if (!ctx || (rule != FSM_RAPIDITY_TAN && rule != FSM_RAPIDITY_GAUSS_HERMITE)) return FSM_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lock(ctx->call_mu);
  ctx->rapidity_rule = rule;
  return FSM_OK;
}

int fsm_gauss_hermite(int n, double *x, double *w) { return gauss_hermite_rule(n, x, w); }

int fsm_set_path(fsm_ctx *ctx, int path) {
  if (!ctx) return FSM_ERR_ARG;
  if (path != FSM_PATH_FP32 && path != FSM_PATH_MIXED) return FSM_ERR_ARG;
  ctx->path = path;
  return FSM_OK;
}

namespace {
// MFMA instructions of one event's MFMA stream stage on tile shape `variant`,
// following the kernel's decisions (fsm_stream2d.hip, stream2d_mfma_tile): a
// wave with no live column runs none; a wave whose cells keep every sample of
// a group runs the group's pair MFMAs for each of its XW columns and its row
// blocks holding live cells; otherwise each offset's MFMAs for those row
// blocks and the columns whose floor indices can be in range.  Tiles that
// fall back to the FP32 loop for spline undershoot are not subtracted.
long mfma_count(const Plan &pl, int variant, int nx, int ny) {
  const int TX = fsm_stream2d_tile_w(variant), TY = fsm_stream2d_tile_h(variant), XW = TX / 4;
  long n = 0;
  for (int x0 = 0; x0 < nx; x0 += TX)
    for (int y0 = 0; y0 < ny; y0 += TY) {
      const int ylast = std::min(y0 + TY - 1, ny - 1);
      const int nyb = y0 + 16 < ny ? 2 : 1;
      for (int w = 0; w < 4; ++w) {
        const int xw0 = x0 + w * XW, xw1 = std::min(xw0 + XW - 1, nx - 1);
        if (xw0 >= nx) continue;
        for (const StencilGroup &G : pl.mgroups_h) {
          const bool interior = (xw0 + G.bx0 >= 1) && (xw1 + G.bx1 - 1 <= nx - 2) && (y0 + G.by0 >= 1) &&
                                (ylast + G.by1 - 1 <= ny - 2);
          if (interior) {
            n += (long)G.pad1 * XW * nyb;
            continue;
          }
          for (int o = G.j0; o < G.j1; ++o) {
            const int ox = pl.moffs_h[o].x, oy = pl.moffs_h[o].y;
            for (int yb = 0; yb < nyb; ++yb) {
              const int ylo = y0 + yb * 16 + oy;
              if (ylo + 15 < 1 || ylo > ny - 2) continue;
              for (int xc = 0; xc < XW; ++xc)
                if ((unsigned)(xw0 + xc + ox - 1) <= (unsigned)(nx - 3)) ++n;
            }
          }
        }
      }
    }
  return n;
}
}  // namespace

int fsm_plan_describe(fsm_ctx *ctx, const fsm_params *p, fsm_plan_info *info) {
  if (!ctx || !p || !info) return FSM_ERR_ARG;
  FSM_TRY(hipSetDevice(ctx->device));
  Plan *pl = nullptr;
  int st = get_plan(ctx, *p, &pl);
  if (st) return st;
  memset(info, 0, sizeof(*info));
  info->dim3 = pl->is3d ? 1 : 0;
  info->samples_kept = pl->samples_kept;
  info->samples_evaluated = pl->samples_evaluated;
  info->est_batch_us = (float)event_cost_us(*pl, *p, ctx->path, false);
  if (pl->is3d) {
    info->tiles = pl->ev3.tiles * pl->ev3.neta;
    info->n_phi = pl->ev3.nphi;
    info->lds_bytes = pl->ev3.lds_floats * (int)sizeof(float);
    return FSM_OK;
  }
  Event2D ev = pl->ev2;
  ev.mfma = ctx->path == FSM_PATH_MIXED;
  info->variant = ev.variant;
  info->tile_w = fsm_stream2d_tile_w(ev.variant);
  info->tile_h = fsm_stream2d_tile_h(ev.variant);
  info->tiles = ev.tiles_x * ev.tiles_y;
  info->n_phi = p->DIM_PHIP;
  info->n_offsets = ev.n_moff;
  info->n_offset_groups = ev.n_mgroups;
  info->n_vector_samples = ev.n_amb;
  info->lds_bytes = (int)fsm_stream2d_event_lds_bytes(ev);
  const int bv = pl->bvariant[ctx->path];
  info->batch_variant = bv;
  info->batch_tiles = ((p->DIM_X + fsm_stream2d_tile_w(bv) - 1) / fsm_stream2d_tile_w(bv)) *
                      ((p->DIM_Y + fsm_stream2d_tile_h(bv) - 1) / fsm_stream2d_tile_h(bv));
  info->tuned = pl->tuned[ctx->path] ? 1 : 0;
  for (int v = 0; v < 3; ++v) info->tune_us[v] = pl->tune_ms[ctx->path][v] * 1000.0f;
  info->mfma_batch = (int)mfma_count(*pl, bv, p->DIM_X, p->DIM_Y);
  info->mfma_one = (int)mfma_count(*pl, ev.variant, p->DIM_X, p->DIM_Y);
  return FSM_OK;
}

int fsm_set_tile_policy(fsm_ctx *ctx, int policy) {
  if (!ctx || (policy != FSM_TILES_AUTO && policy != FSM_TILES_BATCH)) return FSM_ERR_ARG;
  ctx->tile_policy = policy;
  return FSM_OK;
}

int fsm_set_autotune(fsm_ctx *ctx, int on) {
  if (!ctx) return FSM_ERR_ARG;
  ctx->autotune = on ? 1 : 0;
  return FSM_OK;
}

int fsm_autotune(fsm_ctx *ctx, const fsm_params *p, int *variant) {
  if (!ctx || !p) return FSM_ERR_ARG;
  FSM_TRY(hipSetDevice(ctx->device));
  std::shared_ptr<Plan> pl;
  int st = get_plan_shared(ctx, *p, &pl);
  if (st) return st;
  if (!pl->is3d) {
    std::lock_guard<std::mutex> lock(ctx->tune_mu);
    st = autotune2d(ctx, *pl, *p);
    if (st) return st;
  }
  if (variant) *variant = pl->is3d ? 0 : pl->bvariant[ctx->path];
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
  CallOrder order(ctx, (hipStream_t)stream);
  FSM_TRY(order.begin());
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
  hipStream_t s = (hipStream_t)stream;
  CallOrder order(ctx, s);
  FSM_TRY(order.begin());
  std::shared_ptr<Plan> plan;
  int st = get_plan_shared(ctx, *p, &plan);
  if (st) return st;
  Plan *pl = plan.get();
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
  CallOrder order(ctx, s);  // the timing events are the context's too
  FSM_TRY(order.begin());
  EventLandau L = landau_event(*p, d_T, 0);
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

}  // extern "C"

namespace {
// The stage entry points on the reference's host arrays (planes of DIM
// floats, one pointer per component): copy in, one kernel, copy out, on the
// null stream, holding the context's call lock.
struct StageIO {
  fsm_ctx *ctx;
  size_t D;
  float *base = nullptr;
  hipStream_t s = nullptr;
  int put(const float *const *planes, int n, size_t at) {
    for (int c = 0; c < n; ++c) {
      if (!planes[c]) return FSM_ERR_ARG;
      FSM_TRY(hipMemcpyAsync(base + (at + c) * D, planes[c], D * sizeof(float), hipMemcpyHostToDevice, s));
    }
    return FSM_OK;
  }
  int get(float *const *planes, int n, size_t at) {
    for (int c = 0; c < n; ++c) {
      if (!planes[c]) return FSM_ERR_ARG;
      FSM_TRY(hipMemcpyAsync(planes[c], base + (at + c) * D, D * sizeof(float), hipMemcpyDeviceToHost, s));
    }
    return FSM_OK;
  }
};

int stage_begin(fsm_ctx *ctx, const fsm_params *p, size_t planes, StageIO *io) {
  if (!ctx || !p || p->DIM <= 0 || p->DIM != p->DIM_X * p->DIM_Y * p->DIM_ETA) return FSM_ERR_ARG;
  FSM_TRY(hipSetDevice(ctx->device));
  io->ctx = ctx;
  io->D = (size_t)p->DIM;
  FSM_TRY(ctx->d_stage.ensure(planes * io->D));
  io->base = ctx->d_stage.p;
  return FSM_OK;
}

StageCells stage_args(const fsm_params &p, int stage) {
  StageCells a;
  memset(&a, 0, sizeof(a));
  a.dim = p.DIM;
  a.stage = stage;
  a.eos_type = p.EOS_TYPE;
  a.tau = p.TAU;
  const float t2 = p.TAU * p.TAU;  // a float product, as the reference forms it
  a.inv_tau2 = 1.0 / (double)t2;
  return a;
}
}  // namespace

extern "C" {

int fsm_solve_eigen_system(fsm_ctx *ctx, const fsm_params *p, const float *const T[10], float *e,
                           float *const u[4]) {
  if (!T || !e || !u) return FSM_ERR_ARG;
  if (!ctx) return FSM_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lock(ctx->call_mu);  // ctx->d_stage
  StageIO io;
  int st = stage_begin(ctx, p, 15, &io);  // T (10), e, u (4)
  if (st) return st;
  CallOrder order(ctx, io.s);
  FSM_TRY(order.begin());
  if ((st = io.put(T, 10, 0))) return st;
  EventLandau L = landau_event(*p, io.base, 0);
  L.fe = io.base + 10 * io.D;
  L.fu = io.base + 11 * io.D;
  FSM_TRY(fsm_launch_landau(L, io.s));
  float *ep[1] = {e};
  if ((st = io.get(ep, 1, 10)) || (st = io.get(u, 4, 11))) return st;
  FSM_TRY(hipStreamSynchronize(io.s));
  return FSM_OK;
}

int fsm_calculate_pressure(fsm_ctx *ctx, const fsm_params *p, const float *e, const float *nB, float *pr) {
  (void)nB;  // none of the reference's equations of state reads it (src/EquationOfState.cpp:5-80)
  if (!e || !pr) return FSM_ERR_ARG;
  if (!ctx) return FSM_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lock(ctx->call_mu);  // ctx->d_stage
  StageIO io;
  int st = stage_begin(ctx, p, 2, &io);
  if (st) return st;
  CallOrder order(ctx, io.s);
  FSM_TRY(order.begin());
  const float *ein[1] = {e};
  if ((st = io.put(ein, 1, 0))) return st;
  StageCells a = stage_args(*p, FSM_STAGE_PRESSURE);
  a.e = io.base;
  a.p_out = io.base + io.D;
  FSM_TRY(fsm_launch_stage_cells(a, io.s));
  float *pout[1] = {pr};
  if ((st = io.get(pout, 1, 1))) return st;
  FSM_TRY(hipStreamSynchronize(io.s));
  return FSM_OK;
}

int fsm_calculate_bulk_pressure(fsm_ctx *ctx, const fsm_params *p, const float *const T[10], const float *e,
                                const float *pr, float *Pi) {
  if (!T || !e || !pr || !Pi) return FSM_ERR_ARG;
  if (!ctx) return FSM_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lock(ctx->call_mu);  // ctx->d_stage
  StageIO io;
  int st = stage_begin(ctx, p, 13, &io);  // T (10), e, p, Pi
  if (st) return st;
  CallOrder order(ctx, io.s);
  FSM_TRY(order.begin());
  const float *in2[2] = {e, pr};
  if ((st = io.put(T, 10, 0)) || (st = io.put(in2, 2, 10))) return st;
  StageCells a = stage_args(*p, FSM_STAGE_BULK);
  a.T = io.base;
  a.e = io.base + 10 * io.D;
  a.p = io.base + 11 * io.D;
  a.Pi_out = io.base + 12 * io.D;
  FSM_TRY(fsm_launch_stage_cells(a, io.s));
  float *out[1] = {Pi};
  if ((st = io.get(out, 1, 12))) return st;
  FSM_TRY(hipStreamSynchronize(io.s));
  return FSM_OK;
}

int fsm_calculate_shear_visc_tensor(fsm_ctx *ctx, const fsm_params *p, const float *const T[10], const float *e,
                                    const float *const u[4], const float *pr, const float *Pi, float *const pi[10]) {
  if (!T || !e || !u || !pr || !Pi || !pi) return FSM_ERR_ARG;
  if (!ctx) return FSM_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lock(ctx->call_mu);  // ctx->d_stage
  StageIO io;
  int st = stage_begin(ctx, p, 27, &io);  // T (10), e, u (4), p, Pi, pi (10)
  if (st) return st;
  CallOrder order(ctx, io.s);
  FSM_TRY(order.begin());
  const float *ein[1] = {e}, *pin[2] = {pr, Pi};
  if ((st = io.put(T, 10, 0)) || (st = io.put(ein, 1, 10)) || (st = io.put(u, 4, 11)) || (st = io.put(pin, 2, 15)))
    return st;
  StageCells a = stage_args(*p, FSM_STAGE_SHEAR);
  a.T = io.base;
  a.e = io.base + 10 * io.D;
  a.u = io.base + 11 * io.D;
  a.p = io.base + 15 * io.D;
  a.Pi = io.base + 16 * io.D;
  a.pi_out = io.base + 17 * io.D;
  FSM_TRY(fsm_launch_stage_cells(a, io.s));
  if ((st = io.get(pi, 10, 17))) return st;
  FSM_TRY(hipStreamSynchronize(io.s));
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
  std::lock_guard<std::recursive_mutex> lock(ctx->call_mu);  // ctx->d_in, d_out
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

int fsm_run_baryon_device(fsm_ctx *ctx, const fsm_params *p, const float *d_e_GeV, const float *d_nB,
                          double *d_out, void *stream) {
  if (!ctx || !p || !d_e_GeV || !d_out) return FSM_ERR_ARG;
  if (!p->BARYON || !d_nB) return fsm_run_device(ctx, p, d_e_GeV, d_out, stream);
  int st = check_supported(*p);
  if (st) return st;
  FSM_TRY(hipSetDevice(ctx->device));
  CallOrder order(ctx, (hipStream_t)stream);
  FSM_TRY(order.begin());
  float *Tscr = nullptr, *Jscr = nullptr;
  if (p->DIM_ETA == 1) {
    FSM_TRY(ctx->d_T.ensure((size_t)10 * p->DIM));
    FSM_TRY(ctx->d_J.ensure((size_t)4 * p->DIM));
    Tscr = ctx->d_T.p;
    Jscr = ctx->d_J.p;
  }
  return run_device_impl(ctx, p, d_e_GeV, d_out, Tscr, (hipStream_t)stream, d_nB, Jscr);
}

int fsm_run_baryon(fsm_ctx *ctx, fsm_params *p, const float *e_GeV, const float *nB_in, size_t n,
                   double *const out[FSM_NUM_OUTPUTS], double *nB_out, double *const V_out[4],
                   float *tau_landau_match) {
  if (!ctx || !p || !e_GeV || !out) return FSM_ERR_ARG;
  if (!p->BARYON) return fsm_run(ctx, p, e_GeV, n, out, tau_landau_match);
  if (!nB_in) return FSM_ERR_ARG;
  for (int k = 0; k < FSM_NUM_OUTPUTS; ++k)
    if (!out[k]) return FSM_ERR_ARG;
  int st = fsm_params_finalize(p, e_GeV, n);
  if (st) return st;
  if (tau_landau_match) *tau_landau_match = p->TAU;
  FSM_TRY(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lock(ctx->call_mu);  // ctx->d_in, d_nb, d_out
  const size_t D = (size_t)p->DIM;
  FSM_TRY(ctx->d_in.ensure(D));
  FSM_TRY(ctx->d_nb.ensure(D));
  FSM_TRY(ctx->d_out.ensure(D * FSM_NUM_OUTPUTS_BARYON));
  hipStream_t s = nullptr;
  FSM_TRY(hipMemcpyAsync(ctx->d_in.p, e_GeV, D * sizeof(float), hipMemcpyHostToDevice, s));
  FSM_TRY(hipMemcpyAsync(ctx->d_nb.p, nB_in, D * sizeof(float), hipMemcpyHostToDevice, s));
  st = fsm_run_baryon_device(ctx, p, ctx->d_in.p, ctx->d_nb.p, ctx->d_out.p, s);
  if (st) return st;
  for (int k = 0; k < FSM_NUM_OUTPUTS; ++k)
    FSM_TRY(hipMemcpyAsync(out[k], ctx->d_out.p + k * D, D * sizeof(double), hipMemcpyDeviceToHost, s));
  if (nB_out)
    FSM_TRY(hipMemcpyAsync(nB_out, ctx->d_out.p + FSM_NUM_OUTPUTS * D, D * sizeof(double), hipMemcpyDeviceToHost, s));
  if (V_out)
    for (int c = 0; c < 4; ++c)
      if (V_out[c])
        FSM_TRY(hipMemcpyAsync(V_out[c], ctx->d_out.p + (FSM_NUM_OUTPUTS + 1 + c) * D, D * sizeof(double),
                               hipMemcpyDeviceToHost, s));
  FSM_TRY(hipStreamSynchronize(s));
  return FSM_OK;
}

int fsm_diagnostics_device(fsm_ctx *ctx, const fsm_params *p, const float *d_e, const float *d_u, const float *d_p,
                           const float *d_Pi, const float *d_pi, const fsm_diag_out *out, void *stream) {
  if (!ctx || !p || !out) return FSM_ERR_ARG;
  bool pdu = false, bw = false;
  for (int c = 0; c < 4; ++c) pdu |= out->pi_dot_u[c] != nullptr;
  for (int c = 0; c < 4; ++c) bw |= out->beta[c] != nullptr;
  for (int c = 0; c < 6; ++c) bw |= out->omega[c] != nullptr;
  // every requested output needs its inputs
  if ((out->R_Pi_Inv && (!d_p || !d_Pi)) || (out->R_pimunu_Inv && (!d_pi || !d_e || !d_p)) ||
      (pdu && (!d_u || !d_pi)) || (out->pi_mu_mu && !d_pi) || (bw && (!d_e || !d_u)))
    return FSM_ERR_ARG;
  FSM_TRY(hipSetDevice(ctx->device));
  FSM_TRY(fsm_launch_diagnostics(*p, d_e, d_u, d_p, d_Pi, d_pi, *out, (hipStream_t)stream));
  return FSM_OK;
}

int fsm_run_files(fsm_ctx *ctx, fsm_params *p, const float *e_in, const float *nB_in, size_t n,
                  double *const out[FSM_NUM_OUTPUTS], const char *output_dir, fsm_run_report *rep) {
#pragma clang fp contract(off)
  if (!ctx || !p || !e_in) return FSM_ERR_ARG;
  const bool baryon = p->BARYON != 0;
  if (baryon && !nB_in) return FSM_ERR_ARG;
  int st = fsm_params_finalize(p, e_in, n);
  if (st) return st;
  const size_t D = (size_t)p->DIM;
  const float hbarc = 0.197326938f;
  // the density the reference streams, fm^-4 (src/FreestreamMilne.cpp:267 for the vector input)
  std::vector<float> e0(D);
  for (size_t is = 0; is < D; ++is) e0[is] = p->IC_ENERGY == 5 ? e_in[is] * 1.0f / hbarc + 1.0e-7f : e_in[is];
  // the step itself, with the values filled whatever OUTPUTFORMAT is (the
  // reference writes its files from its float arrays in every format)
  fsm_params q = *p;
  q.OUTPUTFORMAT = 2;
  const int np = baryon ? FSM_NUM_OUTPUTS_BARYON : FSM_NUM_OUTPUTS;
  FSM_TRY(hipSetDevice(ctx->device));
  hipStream_t s = nullptr;
  DevBuf<float> d_in, d_nb, d_T, d_f;
  DevBuf<double> d_out;
  FSM_TRY(d_in.ensure(D));
  FSM_TRY(d_out.ensure(D * np));
  FSM_TRY(d_T.ensure(10 * D));
  FSM_TRY(hipMemcpyAsync(d_in.p, e_in, D * sizeof(float), hipMemcpyHostToDevice, s));
  if (baryon) {
    FSM_TRY(d_nb.ensure(D));
    FSM_TRY(hipMemcpyAsync(d_nb.p, nB_in, D * sizeof(float), hipMemcpyHostToDevice, s));
    st = fsm_run_baryon_device(ctx, &q, d_in.p, d_nb.p, d_out.p, s);
  } else {
    st = fsm_run_device(ctx, &q, d_in.p, d_out.p, s);
  }
  if (st) return st;
  st = fsm_stream_tmunu_device(ctx, &q, d_in.p, d_T.p, s);
  if (st) return st;
  std::vector<double> h((size_t)np * D);
  std::vector<float> T00(D);
  FSM_TRY(hipMemcpyAsync(h.data(), d_out.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  FSM_TRY(hipMemcpyAsync(T00.data(), d_T.p, D * sizeof(float), hipMemcpyDeviceToHost, s));
  FSM_TRY(hipStreamSynchronize(s));
  // float fields back from the doubles: (double)x * hbarc is exact, so the
  // division returns x; e, p, u[4], pi[10], Pi (+ n_B, V[4]) in planes f[0..]
  std::vector<float> f((size_t)np * D);
  for (int c = 0; c < np; ++c) {
    const bool scaled = c < 2 || (c >= 6 && c < FSM_NUM_OUTPUTS);
    for (size_t is = 0; is < D; ++is)
      f[c * D + is] = scaled ? (float)(h[c * D + is] / (double)hbarc) : (float)h[c * D + is];
  }
  const float *fe = &f[0], *fp = &f[D], *fu = &f[2 * D], *fpi = &f[6 * D], *fPi = &f[16 * D];
  // diagnostics on the device from the float fields
  const int nd = 1 + 1 + 4 + 1 + 4 + 6;
  FSM_TRY(d_f.ensure(17 * D + nd * D));
  FSM_TRY(hipMemcpyAsync(d_f.p, f.data(), 17 * D * sizeof(float), hipMemcpyHostToDevice, s));
  float *dd = d_f.p + 17 * D;
  fsm_diag_out o;
  o.R_Pi_Inv = dd;
  o.R_pimunu_Inv = dd + D;
  for (int c = 0; c < 4; ++c) o.pi_dot_u[c] = dd + (2 + c) * D;
  o.pi_mu_mu = dd + 6 * D;
  for (int c = 0; c < 4; ++c) o.beta[c] = dd + (7 + c) * D;
  for (int c = 0; c < 6; ++c) o.omega[c] = dd + (11 + c) * D;
  st = fsm_diagnostics_device(ctx, &q, d_f.p, d_f.p + 2 * D, d_f.p + D, d_f.p + 16 * D, d_f.p + 6 * D, &o, s);
  if (st) return st;
  std::vector<float> diag((size_t)nd * D);
  FSM_TRY(hipMemcpyAsync(diag.data(), dd, diag.size() * sizeof(float), hipMemcpyDeviceToHost, s));
  FSM_TRY(hipStreamSynchronize(s));
  // report (the values the reference prints)
  float e_before = 0.0f;
  for (size_t is = 0; is < D; ++is) e_before += e0[is];
  if (p->DIM_ETA > 1)
    e_before *= (p->TAU0 * p->DX * p->DY * p->DETA);
  else
    e_before *= (p->DX * p->DY);
  float total = 0.0f, inside = 0.0f;
  fsm_energy_sums(p, T00.data(), fe, &total, &inside);
  const int NX = p->DIM_X, NY = p->DIM_Y, NE = p->DIM_ETA;
  const size_t ctr = (size_t)(NY * NE) * ((NX - 1) / 2) + (size_t)NE * ((NY - 1) / 2) + (NE - 1) / 2;
  if (rep) {
    memset(rep, 0, sizeof(*rep));
    rep->total_energy_before = e_before;
    rep->total_energy_after = total;
    rep->energy_inside = inside;
    rep->R_Pi_Inv_center = diag[ctr];
    rep->R_pimunu_Inv_center = diag[D + ctr];
    rep->tau_landau_match = p->TAU;
    rep->dtau = p->DTAU;
    rep->dim_phip = p->DIM_PHIP;
  }
  // host vectors: the 17 outputs, zero unless OUTPUTFORMAT is 2 or 3 (src/FreestreamMilne.cpp:695)
  if (out) {
    const bool fill = p->OUTPUTFORMAT == 2 || p->OUTPUTFORMAT == 3;
    for (int c = 0; c < FSM_NUM_OUTPUTS; ++c)
      if (out[c]) {
        if (fill)
          memcpy(out[c], &h[c * D], D * sizeof(double));
        else
          memset(out[c], 0, D * sizeof(double));
      }
  }
  // the reference's output/*.dat files, in its order (src/FreestreamMilne.cpp:310-313,
  // :555-559, :565-673); a missing directory writes nothing, as the
  // reference's unopened streams do
  if (output_dir) {
    struct stat sb;
    if (stat(output_dir, &sb) != 0 || !S_ISDIR(sb.st_mode)) return FSM_OK;
    auto path = [&](const char *name) {
      std::string s = output_dir;
      if (!s.empty() && s.back() != '/') s += '/';
      return s + name + ".dat";
    };
    auto scalar = [&](const char *name, const float *v) { fsm_write_scalar(p, path(name).c_str(), v); };
    auto proj = [&](const char *name, const float *v) { fsm_write_projection(p, path(name).c_str(), v); };
    std::vector<float> nb0;
    if (baryon) nb0.assign(nB_in, nB_in + D);
    scalar("initial_e", e0.data());
    if (baryon) scalar("initial_nB", nb0.data());
    proj("initial_e_projection", e0.data());
    if (baryon) proj("initial_nB_projection", nb0.data());
    fsm_write_energy_inside(path("energy_inside_switch_surf").c_str(), inside);
    proj("R_Pi_Inv_projection", &diag[0]);
    proj("R_pimunu_Inv_projection", &diag[D]);
    static const char *pdu[4] = {"pi_dot_u_tau_projection", "pi_dot_u_x_projection", "pi_dot_u_y_projection",
                                 "pi_dot_u_eta_projection"};
    for (int c = 0; c < 4; ++c) proj(pdu[c], &diag[(2 + c) * D]);
    proj("pi_mu_mu_projection", &diag[6 * D]);
    scalar("e", fe);
    scalar("p", fp);
    scalar("bulk_PI", fPi);
    proj("e_projection", fe);
    proj("p_projection", fp);
    proj("bulk_PI_projection", fPi);
    static const char *un[4] = {"ut", "ux", "uy", "un"};
    for (int c = 0; c < 4; ++c) scalar(un[c], fu + c * D);
    for (int c = 0; c < 4; ++c) proj((std::string(un[c]) + "_projection").c_str(), fu + c * D);
    static const char *pn[10] = {"pitt", "pitx", "pity", "pitn", "pixx", "pixy", "pixn", "piyy", "piyn", "pinn"};
    for (int c = 0; c < 10; ++c) scalar(pn[c], fpi + c * D);
    for (int c = 0; c < 10; ++c) proj((std::string(pn[c]) + "_projection").c_str(), fpi + c * D);
    proj("beta_x_projection", &diag[8 * D]);
    scalar("beta_x", &diag[8 * D]);
    proj("beta_y_projection", &diag[9 * D]);
    scalar("beta_y", &diag[9 * D]);
    proj("w_xy_projection", &diag[14 * D]);
    scalar("w_xy", &diag[14 * D]);
    if (baryon) {
      const float *fnb = &f[FSM_NUM_OUTPUTS * D], *fV = &f[(FSM_NUM_OUTPUTS + 1) * D];
      scalar("nB", fnb);
      proj("nB_projection", fnb);
      scalar("V_x", fV + D);
      scalar("V_y", fV + 2 * D);
      scalar("V_eta", fV + 3 * D);
    }
  }
  return FSM_OK;
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
  const float *d_nb = nullptr;  // baryon events: initial n_B (charge pass)
  double *d_out;
  char *ws = nullptr;  // caller-provided workspace (event_workspace_bytes), else the batch arena
};

inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Device workspace of one event in a batch: spline planes + T (2-D), or
// source planes + eta-major T (3-D).
size_t spline_flag_count(const fsm_params &p) {
  return (size_t)((p.DIM_X + FSMSP_TX - 1) / FSMSP_TX) * ((p.DIM_Y + FSMSP_TY - 1) / FSMSP_TY);
}

size_t event_workspace_bytes(const fsm_params &p);

// Baryon events stream twice (energy and charge), each pass with its own workspace.
size_t event_workspace_bytes_all(const fsm_params &p, bool baryon) {
  return event_workspace_bytes(p) * (baryon ? 2 : 1);
}

// Row pitch of a batch's internal 2-D T (cells): DIM_Y rounded up to 32.
inline int padded_row(int ny) { return (ny + 31) & ~31; }

size_t event_workspace_bytes(const fsm_params &p) {
  if (p.DIM_ETA == 1)
    return align256((size_t)p.DIM_X * p.DIM_Y * sizeof(float4)) +
           align256((size_t)10 * p.DIM_X * padded_row(p.DIM_Y) * sizeof(float)) +
           align256((size_t)p.DIM_X * p.DIM_Y * sizeof(uint4)) + align256(spline_flag_count(p) * sizeof(unsigned)) +
           align256((FSM_KS_PARTS + 1) * sizeof(float));
  return align256((size_t)p.DIM * sizeof(float)) + align256((size_t)10 * p.DIM * sizeof(float));
}

// Descriptor of one 2-D event of a batch whose workspace starts at `cur`
// (layout of event_workspace_bytes): spline planes, T, MFMA planes, spline
// tile flags, knot scale.  Advances `cur` past the workspace.  pad_rows: T
// rows at the padded pitch (read through EventLandau::t_row), else the
// reference layout.
Event2D event2d_in_workspace(const Plan &pl, const fsm_params &p, const float *d_e, int path, int variant,
                             char *&cur, bool pad_rows) {
  Event2D d = pl.ev2;
  d.e_gev = d_e;
  set_variant(d, variant);
  d.planes = (float4 *)cur;
  cur += align256((size_t)p.DIM_X * p.DIM_Y * sizeof(float4));
  d.T = (float *)cur;
  cur += align256((size_t)10 * p.DIM_X * padded_row(p.DIM_Y) * sizeof(float));
  d.t_row = pad_rows ? padded_row(p.DIM_Y) : p.DIM_Y;
  d.t_plane = (long)p.DIM_X * d.t_row;
  uint4 *ph = (uint4 *)cur;
  cur += align256((size_t)p.DIM_X * p.DIM_Y * sizeof(uint4));
  d.sp_flags = (unsigned *)cur;
  cur += align256(spline_flag_count(p) * sizeof(unsigned));
  d.mscale = (float *)cur;
  cur += align256((FSM_KS_PARTS + 1) * sizeof(float));
  d.mfma = path == FSM_PATH_MIXED ? 1 : 0;
  d.planes_h = d.mfma ? ph : nullptr;
  return d;
}

// Batches: the knot pass writes the knot values z into the first plane of
// the event's T, which the batch spline kernel reads (each input value
// converted once, not once per tile halo that holds it); the stream kernel,
// launched after the spline kernel, overwrites that plane.  The pass also
// writes the partial maxima of |z| that the MFMA path's knot scale needs.
inline void keep_knots_in_t(Event2D &d) { d.zk = d.T; }

// Autotuner of the batch tile shape: times the 2-D stream kernel of this
// event shape over a synthetic batch (a smooth profile, enough events to
// give every CU four tiles of the widest shape) for each tile shape on the
// context's path, and keeps the fastest as the plan's batch shape.
int autotune2d(fsm_ctx *ctx, Plan &pl, const fsm_params &p) {
  const int path = ctx->path;
  const int nx = p.DIM_X, ny = p.DIM_Y;
  const long t0 = (long)((nx + fsm_stream2d_tile_w(0) - 1) / fsm_stream2d_tile_w(0)) *
                  ((ny + fsm_stream2d_tile_h(0) - 1) / fsm_stream2d_tile_h(0));
  const int n = (int)std::min<long>(16, std::max<long>(1, (4L * ctx->num_cus + t0 - 1) / t0));
  std::vector<float> h((size_t)p.DIM);
  for (int ix = 0; ix < nx; ++ix)
    for (int iy = 0; iy < ny; ++iy) {
      const double x = (ix - 0.5 * (nx - 1)) / (0.2 * nx), y = (iy - 0.5 * (ny - 1)) / (0.2 * ny);
      h[(size_t)ix * ny + iy] = (float)(10.0 * exp(-0.5 * (x * x + y * y)));
    }
  DevBuf<float> d_in;
  FSM_TRY(d_in.upload(h));
  const size_t wsb = event_workspace_bytes(p);
  DevBuf<char> ws;
  FSM_TRY(ws.ensure(wsb * n));
  DevBuf<Event2D> d_ev;
  DevBuf<int4> d_items;
  hipStream_t s = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  FSM_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  int st = FSM_OK;
  float best = 1e30f;
  int bestv = pl.bvariant[path];
  if (hipEventCreate(&ev0) != hipSuccess || hipEventCreate(&ev1) != hipSuccess) st = FSM_ERR_DEVICE;
  for (int v = 0; v < 3 && st == FSM_OK; ++v) {
    std::vector<Event2D> evs;
    std::vector<int4> ks, sp, s2;
    size_t lds = 0;
    char *cur = ws.p;
    const int sty = (ny + FSMSP_TY - 1) / FSMSP_TY;
    for (int i = 0; i < n; ++i) {
      Event2D d = event2d_in_workspace(pl, p, d_in.p, path, v, cur, true);
      keep_knots_in_t(d);
      for (int part = 0; part < FSM_KS_PARTS; ++part) ks.push_back(make_int4(i, part, 0, 0));
      for (int x = 0; x < (nx + FSMSP2_TX - 1) / FSMSP2_TX; ++x)
        for (int y = 0; y < sty; ++y) sp.push_back(make_int4(i, x, y, 0));
      for (int y = 0; y < d.tiles_y; ++y)
        for (int x = 0; x < d.tiles_x; ++x) s2.push_back(make_int4(i, x, y, 0));
      lds = std::max(lds, fsm_stream2d_event_lds_bytes(d));
      evs.push_back(d);
    }
    std::vector<int4> items(ks);
    items.insert(items.end(), sp.begin(), sp.end());
    items.insert(items.end(), s2.begin(), s2.end());
    if (d_ev.upload(evs) != hipSuccess || d_items.upload(items) != hipSuccess) {
      st = FSM_ERR_DEVICE;
      break;
    }
    const int4 *it = d_items.p;
    const int nks = (int)ks.size(), nsp = (int)sp.size(), ns2 = (int)s2.size();
    hipError_t e = fsm_launch_knot_scale_batch(d_ev.p, it, nks, s);
    if (e == hipSuccess) e = fsm_launch_spline2d_batch(d_ev.p, it + nks, nsp, s);
    if (e == hipSuccess) e = fsm_launch_stream2d_batch(d_ev.p, it + nks + nsp, ns2, lds, path == FSM_PATH_MIXED, s);
    if (e == hipSuccess) e = hipEventRecord(ev0, s);
    for (int r = 0; r < 3 && e == hipSuccess; ++r)
      e = fsm_launch_stream2d_batch(d_ev.p, it + nks + nsp, ns2, lds, path == FSM_PATH_MIXED, s);
    if (e == hipSuccess) e = hipEventRecord(ev1, s);
    if (e == hipSuccess) e = hipEventSynchronize(ev1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0, ev1);
    if (e != hipSuccess) {
      st = hip_status(e);
      break;
    }
    pl.tune_ms[path][v] = ms / (3.0f * n);
    if (ms < best) {
      best = ms;
      bestv = v;
    }
  }
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  (void)hipStreamSynchronize(s);
  (void)hipStreamDestroy(s);
  if (st == FSM_OK) {
    pl.bvariant[path] = bestv;
    pl.tuned[path] = true;
  }
  return st;
}

// Tunes a 2-D plan's batch shape for the context's path once, if enabled.
int maybe_tune(fsm_ctx *ctx, Plan &pl, const fsm_params &p) {
  if (!ctx->autotune || pl.is3d) return FSM_OK;
  std::lock_guard<std::mutex> lock(ctx->tune_mu);
  if (pl.tuned[ctx->path]) return FSM_OK;
  return autotune2d(ctx, pl, p);
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
  // work-item ranges per stage: knot scale, spline, 2-D stream, 3-D source, 3-D stream, Landau
  int off_ks = 0, n_ks = 0, off_sp = 0, n_sp = 0, off_s2 = 0, n_s2 = 0;
  int off_src = 0, n_src = 0, off_s3 = 0, n_s3 = 0, off_l = 0, n_l = 0, off_l2 = 0, n_l2 = 0;
  size_t lds2 = 0, lds3 = 0, lds_l = 0;  // dynamic LDS of the stream and Landau launches
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // stream-ordered uploads (used by the server): descriptors go through
  // pinned staging with hipMemcpyAsync on the launch stream
  bool async_upload = false;
  PinnedBuf staging;
  // completion of the last fsm_batch_run: rebuilding the descriptors or the
  // graph waits for it, since that run may still read them on its stream
  hipEvent_t last_run = nullptr;
  bool last_run_recorded = false;

  void wait_last_run() {
    if (last_run_recorded) (void)hipEventSynchronize(last_run);
    last_run_recorded = false;
  }
  void drop_graph() {
    wait_last_run();
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~fsm_batch() {
    drop_graph();
    if (last_run) (void)hipEventDestroy(last_run);
  }
};

namespace {

// Cells per Landau work item of a batch: one per thread of one workgroup.
// The stage is bound by its 17 double planes of writes; with one round per
// workgroup the workgroups running on an XCD at a time write one contiguous
// stretch of each plane (consecutive items), where four rounds per item
// spread them over four times the address range (C1, 64 events: 0.156 ms
// against 0.195 ms with 1024-cell items, the same bits).
const int kLandauChunk = 256;

int batch_build(fsm_batch *b, hipStream_t s) {
  b->drop_graph();
  b->built_path = b->ctx->path;
  std::vector<Event2D> e2;
  std::vector<Event3D> e3;
  std::vector<EventLandau> el;
  std::vector<int4> ks, sp, s2, src, s3, ld, ld2;
  std::vector<float> s3w;  // work estimate of each 3-D stream item
  size_t bytes = 0;
  for (auto &ev : b->evs)
    if (!ev.ws) bytes += event_workspace_bytes_all(ev.p, ev.d_nb != nullptr);
  if (bytes) FSM_TRY(b->arena.ensure(bytes));
  char *arena_cur = b->arena.p;
  b->lds2 = b->lds3 = b->lds_l = 0;
  const int path = b->ctx->path;
  // stage-A work items of one streaming pass (energy, or the charge pass of a
  // baryon event) in the workspace at `cur`; returns its T (or J) and layout
  auto add_pass = [&](const BatchEvent &ev, const float *d_in, bool charge, char *&cur, int *eta_major) {
    const fsm_params &p = ev.p;
    if (!ev.plan->is3d) {
      Event2D d = event2d_in_workspace(*ev.plan, p, d_in, path, ev.plan->bvariant[path], cur, true);
      // the batch's T is read only by its 3x3 Landau launch (k_landau2d_batch)
      d.skip_eta_planes = charge ? 0 : 1;
      if (charge) {
        d.charge = 1;
        d.out_j = 1;
        d.out_scale = 1.0f;  // DIM_RAP = 1
      }
      const int i = (int)e2.size();
      keep_knots_in_t(d);
      for (int part = 0; part < FSM_KS_PARTS; ++part) ks.push_back(make_int4(i, part, 0, 0));
      e2.push_back(d);
      const int sty = (p.DIM_Y + FSMSP_TY - 1) / FSMSP_TY;
      for (int x = 0; x < (p.DIM_X + FSMSP2_TX - 1) / FSMSP2_TX; ++x)
        for (int y = 0; y < sty; ++y) sp.push_back(make_int4(i, x, y, 0));
      for (int y = 0; y < d.tiles_y; ++y)
        for (int x = 0; x < d.tiles_x; ++x) s2.push_back(make_int4(i, x, y, 0));
      b->lds2 = std::max(b->lds2, fsm_stream2d_event_lds_bytes(d));
      *eta_major = 0;
      return (const float *)d.T;
    }
    Event3D d = charge ? ev.plan->ev3b : ev.plan->ev3;
    d.e_gev = d_in;
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
      for (int t = 0; t < d.tiles; ++t) {
        s3.push_back(make_int4(i, t, ie, 0));
        s3w.push_back(ev.plan->eta_work[ie]);
      }
    b->lds3 = std::max(b->lds3, (size_t)d.lds_floats * sizeof(float));
    *eta_major = 1;
    return (const float *)d.T;
  };
  for (auto &ev : b->evs) {
    const fsm_params &p = ev.p;
    const int li = (int)el.size();
    char *cur = ev.ws;
    if (!cur) {
      cur = arena_cur;
      arena_cur += event_workspace_bytes_all(p, ev.d_nb != nullptr);
    }
    int em = 0, jm = 0;
    const float *T = add_pass(ev, ev.d_e, false, cur, &em);
    if (!em && !ev.d_nb) {
      // a 2-D event without a baryon current: the Landau stage runs in its
      // stream tiles' epilogue (no Landau items)
      fuse_landau2d(e2.back(), p, ev.d_out);
      continue;
    }
    EventLandau L = landau_event(p, T, em);
    L.eta_zero = p.DIM_ETA == 1;  // the 2-D stream kernel writes the eta components as zeros
    if (!em) {
      // the batch's 2-D T (and J) rows sit at the padded pitch
      L.t_ny = p.DIM_Y;
      L.t_row = padded_row(p.DIM_Y);
      L.t_plane = (long)p.DIM_X * L.t_row;
    }
    L.out = ev.d_out;
    if (ev.d_nb) {
      L.J = add_pass(ev, ev.d_nb, true, cur, &jm);
      L.out_b = ev.d_out + (size_t)FSM_NUM_OUTPUTS * p.DIM;
    }
    el.push_back(L);
    if (em) {
      // 3-D: eta x transverse tiles of the LDS-transposed Landau form
      const int nxy = p.DIM_X * p.DIM_Y;
      // eta tiles fastest, as in the one-event launch (contiguous output stretches)
      for (int xt = 0; xt < (nxy + FSML_TX - 1) / FSML_TX; ++xt)
        for (int et = 0; et < (p.DIM_ETA + FSML_TE - 1) / FSML_TE; ++et) ld.push_back(make_int4(li, xt, et, 1));
      b->lds_l = std::max(b->lds_l, fsm_landau_lds_bytes(ev.d_nb != nullptr));
    } else {
      // 2-D (DIM_ETA = 1, eta components zero): cell chunks of the 3x3 form
      for (int c = 0; c < p.DIM; c += kLandauChunk) ld2.push_back(make_int4(li, c, std::min(kLandauChunk, p.DIM - c), 0));
    }
  }
  {
    // 3-D stream items by decreasing work (every event's heavy eta planes
    // first), so the tail of the launch is made of short items
    std::vector<int> idx(s3.size());
    for (size_t k = 0; k < idx.size(); ++k) idx[k] = (int)k;
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return s3w[x] > s3w[y]; });
    std::vector<int4> sorted(s3.size());
    for (size_t k = 0; k < idx.size(); ++k) sorted[k] = s3[idx[k]];
    s3.swap(sorted);
  }
  std::vector<int4> items;
  auto put = [&](std::vector<int4> &v, int *off, int *n) {
    *off = (int)items.size();
    *n = (int)v.size();
    items.insert(items.end(), v.begin(), v.end());
  };
  put(ks, &b->off_ks, &b->n_ks);
  put(sp, &b->off_sp, &b->n_sp);
  put(s2, &b->off_s2, &b->n_s2);
  put(src, &b->off_src, &b->n_src);
  put(s3, &b->off_s3, &b->n_s3);
  put(ld, &b->off_l, &b->n_l);
  put(ld2, &b->off_l2, &b->n_l2);
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

int batch_run_graph_or_direct(fsm_batch *b, hipStream_t s);

int batch_launch(fsm_batch *b, hipStream_t s, bool timed) {
  fsm_ctx *ctx = b->ctx;
  const int4 *it = b->d_items.p;
  if (timed) tmark(ctx, 0, s);
  FSM_TRY(fsm_launch_knot_scale_batch(b->d_ev2.p, it + b->off_ks, b->n_ks, s));
  FSM_TRY(fsm_launch_spline2d_batch(b->d_ev2.p, it + b->off_sp, b->n_sp, s));
  FSM_TRY(fsm_launch_src3d_batch(b->d_ev3.p, it + b->off_src, b->n_src, s));
  if (timed) tmark(ctx, 1, s);
  const int mfma = b->built_path == FSM_PATH_MIXED;
  FSM_TRY(fsm_launch_stream2d_batch(b->d_ev2.p, it + b->off_s2, b->n_s2, b->lds2, mfma, s));
  FSM_TRY(fsm_launch_stream3d_batch(b->d_ev3.p, it + b->off_s3, b->n_s3, b->lds3, s));
  if (timed) tmark(ctx, 2, s);
  FSM_TRY(fsm_launch_landau_batch(b->d_evl.p, it + b->off_l, b->n_l, b->lds_l, s));
  FSM_TRY(fsm_launch_landau2d_batch(b->d_evl.p, it + b->off_l2, b->n_l2, s));
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
  st = maybe_tune(b->ctx, *ev.plan, *p);
  if (st) return st;
  ev.p = *p;
  ev.d_e = d_e_GeV;
  ev.d_out = d_out;
  b->evs.push_back(ev);
  b->dirty = true;
  return (int)b->evs.size() - 1;
}

int fsm_batch_add_baryon(fsm_batch *b, const fsm_params *p, const float *d_e_GeV, const float *d_nB,
                         double *d_out) {
  if (!b || !p || !d_nB) return FSM_ERR_ARG;
  if (!p->BARYON) return FSM_ERR_ARG;
  int st = check_supported(*p);
  if (st) return st;
  int i = fsm_batch_add(b, p, d_e_GeV, d_out);
  if (i >= 0) b->evs[i].d_nb = d_nB;
  return i;
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
  if (!b->last_run) FSM_TRY(hipEventCreateWithFlags(&b->last_run, hipEventDisableTiming));
  if (b->dirty || b->built_path != b->ctx->path) {
    int st = batch_build(b, s);
    if (st) return st;
  }
  int st = batch_run_graph_or_direct(b, s);
  if (st) return st;
  FSM_TRY(hipEventRecord(b->last_run, s));
  b->last_run_recorded = true;
  return FSM_OK;
}

}  // extern "C"

namespace {
int batch_run_graph_or_direct(fsm_batch *b, hipStream_t s) {
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
}  // namespace


namespace {
// The cached plan of p's shape, or null: no table is built and the cache is
// not reordered (both could allocate).
std::shared_ptr<Plan> find_plan(fsm_ctx *ctx, const fsm_params &p) {
  std::lock_guard<std::mutex> lock(ctx->plan_mu);
  for (const auto &pl : ctx->plans)
    if (plan_matches(*pl, p, ctx)) return pl;
  return nullptr;
}
}  // namespace

extern "C" {

int fsm_run_device_noalloc(fsm_ctx *ctx, const fsm_params *p, const float *d_e_GeV, double *d_out, void *stream) {
  if (!ctx || !p || !d_e_GeV || !d_out) return FSM_ERR_ARG;
  FSM_TRY(hipSetDevice(ctx->device));
  if (!ctx->noalloc_scratch) return FSM_ERR_DEVICE;
  if (event_workspace_bytes(*p) > FSM_NOALLOC_SCRATCH_BYTES) return FSM_ERR_TOO_LARGE;
  hipStream_t s = (hipStream_t)stream;
  CallOrder order(ctx, s);
  FSM_TRY(order.begin());
  std::shared_ptr<Plan> plan = find_plan(ctx, *p);
  if (!plan) return FSM_ERR_NO_PLAN;
  const Plan &pl = *plan;
  char *ws = (char *)ctx->noalloc_scratch;
  EventLandau L;
  if (!pl.is3d) {
    // the one-event kernels of fsm_run_device, with the spline planes, the
    // fp16 planes, the tile flags, the knot scale and T in the workspace
    const int v = ctx->tile_policy == FSM_TILES_BATCH ? pl.bvariant[ctx->path] : pl.ev2.variant;
    char *cur = ws;
    Event2D ev = event2d_in_workspace(pl, *p, d_e_GeV, ctx->path, v, cur, false);
    fuse_landau2d(ev, *p, d_out);  // the Landau stage in the stream kernel's epilogue
    FSM_TRY(fsm_launch_spline2d(ev, s));
    FSM_TRY(fsm_launch_stream2d(ev, s));
    return FSM_OK;
  } else {
    // source planes and eta-major T in the workspace
    Event3D ev = pl.ev3;
    ev.e_gev = d_e_GeV;
    ev.src = (float *)ws;
    ev.T = (float *)(ws + align256((size_t)p->DIM * sizeof(float)));
    ev.t_plane = p->DIM;
    FSM_TRY(fsm_launch_src3d(ev, s));
    FSM_TRY(fsm_launch_stream3d(ev, s));
    L = landau_event(*p, ev.T, 1);
  }
  L.out = d_out;
  FSM_TRY(fsm_launch_landau(L, s));
  return FSM_OK;
}

}  // extern "C"

// ---- internal entry points for the server (fsm_host.h) ----------------------
namespace fsm_host {

size_t workspace_bytes(const fsm_params &p, bool baryon) { return event_workspace_bytes_all(p, baryon); }

int batch_add_ws(fsm_batch *b, const fsm_params *p, const float *d_e, const float *d_nb, double *d_out, void *ws) {
  int i = d_nb ? fsm_batch_add_baryon(b, p, d_e, d_nb, d_out) : fsm_batch_add(b, p, d_e, d_out);
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

int plan_check(fsm_ctx *ctx, const fsm_params *p, bool baryon, float *cost_us) {
  std::shared_ptr<Plan> sp;
  int st = get_plan_shared(ctx, *p, &sp);
  if (st) return st;
  st = maybe_tune(ctx, *sp, *p);
  if (st) return st;
  if (cost_us) *cost_us = (float)event_cost_us(*sp, *p, ctx->path, baryon);
  return FSM_OK;
}

}  // namespace fsm_host
