// fsm_kernels.hip -- CDNA4 (gfx950) kernels of the free-streaming +
// Landau-matching step.
//
//   k_spline2d   ingest (src/FreestreamMilne.cpp:267) + convertInitialDensity
//                in 2-D (src/FreeStream.cpp:308-322) + the knot derivatives
//                of the natural bicubic spline (src/FreeStream.cpp:121-153),
//                one LDS tile per work item.                       HBM-bound
//   k_stream2d   freeStream 2-D (src/FreeStream.cpp:156-248) fused with
//                calculateHypertrigTable / calculateStressTensor
//                (src/LandauMatch.cpp:14-108): per azimuth sample, the
//                bicubic value (bilinear when the spline is <= 0) of the knot
//                planes staged in LDS, accumulated into six angular moments.
//                                                              FP32 VALU-bound
//   k_src3d      ingest + convertInitialDensity norm in 3-D, transposed to
//                eta-major planes.                                 HBM-bound
//   k_stream3d   freeStream 3-D (src/FreeStream.cpp:190-273) + quadrature:
//                per (eta, rapidity slot) the eta-blended source plane is
//                staged in LDS and each azimuth sample is a bilinear gather.
//                                                              FP32 VALU-bound
//   k_landau     solveEigenSystem (src/LandauMatch.cpp:145-331) by FP64
//                hyperbolic Jacobi rotations, calculatePressure
//                (src/EquationOfState.cpp:5-80), calculateBulkPressure and
//                calculateShearViscTensor (src/LandauMatch.cpp:332-367) and
//                the output pack (src/FreestreamMilne.cpp:695-718).  HBM-bound
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include "fsm_internal.h"
#include "fsm_kernels.h"

namespace {

__device__ __forceinline__ float ingest_e0(float v) {
  // src/FreestreamMilne.cpp:267: v * rescale / (float)hbarc + lower_tolerance
  return v * 1.0f / FSM_HBARC_F + 1.0e-7f;
}

// Knot value of the 2-D spline for one input value: the energy path's
// e0 / (2 pi) (src/FreeStream.cpp:318) or the charge path's n_B * norm
// (src/FreeStream.cpp:384 with a rapidity factor of 1 at eta = rap = 0; the
// sum over rapidity copies is folded into the output scale).
__device__ __forceinline__ float knot_z(const Event2D &ev, float v) {
  if (ev.charge) return v * ev.in_norm;
  return (float)((double)(ev.raw_e ? v : ingest_e0(v)) / (2.0 * M_PI));
}

// ======================================================================
// 2-D spline knot planes
// ======================================================================
#define SP_RX (FSMSP_TX + 2 * FSM_SPLINE_BAND)
#define SP_RY (FSMSP_TY + 2 * FSM_SPLINE_BAND)

// Shared memory of one spline tile: z over the tile plus a band-wide halo,
// z_x over the tile rows plus the y halo, and the band weights of the tile's
// x rows and y columns.
struct SplineLds {
  float zr[SP_RX * SP_RY];
  float zxr[FSMSP_TX * SP_RY];
  float wx[FSMSP_TX * FSM_SPLINE_W];
  float wy[FSMSP_TY * FSM_SPLINE_W];
  float red[4];
  float sc;  // knot scale of the MFMA planes
};

// Halves of the MFMA knot format: h = fp16(x), l = fp16(x - h).
__device__ __forceinline__ void split_half(float x, unsigned short &h, unsigned short &l) {
  const _Float16 hv = (_Float16)x;
  const _Float16 lv = (_Float16)(x - (float)hv);
  h = __builtin_bit_cast(unsigned short, hv);
  l = __builtin_bit_cast(unsigned short, lv);
}

// Smallest Bernstein control point over the four bicubic patches a knot is a
// corner of.  A patch with all 16 control points > 0 is > 0 everywhere
// (convex hull), so the reference never falls back to bilinear inside it.
__device__ __forceinline__ float knot_cp_min(float z, float zy, float zx, float zxy) {
  const float a = zx * (1.0f / 3.0f), b = zy * (1.0f / 3.0f), c = zxy * (1.0f / 9.0f);
  float m = fminf(z, fminf(z + a, z - a));
  m = fminf(m, fminf(z + b, z - b));
  m = fminf(m, fminf(z + a + b + c, z - a - b + c));
  m = fminf(m, fminf(z + a - b - c, z - a + b - c));
  return m;
}

__device__ float knot_scale_from_parts(const Event2D &ev);  // below, with k_knot_scale

__device__ void spline2d_tile(const Event2D &ev, int tx, int ty, SplineLds &L) {
  const int nx = ev.nx, ny = ev.ny;
  const int x0 = tx * FSMSP_TX, y0 = ty * FSMSP_TY;
  const int B = FSM_SPLINE_BAND;
  __syncthreads();
  float zabs = 0.0f;
  for (int k = threadIdx.x; k < SP_RX * SP_RY; k += blockDim.x) {
    int lx = k / SP_RY, ly = k - lx * SP_RY;
    int gx = x0 - B + lx, gy = y0 - B + ly;
    float z = 0.0f;
    if (gx >= 0 && gx < nx && gy >= 0 && gy < ny) z = knot_z(ev, ev.e_gev[(long)gx * ny + gy]);
    L.zr[k] = z;
    zabs = fmaxf(zabs, fabsf(z));
  }
  if (ev.planes_h) {
    // largest |z| within the band of the tile: the scale of the rounding
    // noise in the knot slopes, below which a negative control point is noise
    for (int o = 32; o > 0; o >>= 1) zabs = fmaxf(zabs, __shfl_xor(zabs, o));
    if ((threadIdx.x & 63) == 0) L.red[threadIdx.x >> 6] = zabs;
    if (threadIdx.x < 64) {
      const float sc = knot_scale_from_parts(ev);
      if (threadIdx.x == 0) {
        L.sc = sc;
        ev.mscale[FSM_KS_PARTS] = sc;  // the same value from every block
      }
    }
  }
  for (int k = threadIdx.x; k < FSMSP_TX * FSM_SPLINE_W; k += blockDim.x) {
    const int gx = x0 + k / FSM_SPLINE_W;
    L.wx[k] = gx < nx ? ev.gx[(long)x0 * FSM_SPLINE_W + k] : 0.0f;
  }
  for (int k = threadIdx.x; k < FSMSP_TY * FSM_SPLINE_W; k += blockDim.x) {
    const int gy = y0 + k / FSM_SPLINE_W;
    L.wy[k] = gy < ny ? ev.gy[(long)y0 * FSM_SPLINE_W + k] : 0.0f;
  }
  __syncthreads();
  // z_x (in grid units) for the tile rows and the y-halo.
  for (int k = threadIdx.x; k < FSMSP_TX * SP_RY; k += blockDim.x) {
    int lx = k / SP_RY, ly = k - lx * SP_RY;
    int gx = x0 + lx, gy = y0 - B + ly;
    float acc = 0.0f;
    if (gx < nx && gy >= 0 && gy < ny) {
      const float *w = L.wx + lx * FSM_SPLINE_W;
#pragma unroll
      for (int m = 0; m < FSM_SPLINE_W; ++m) acc = fmaf(w[m], L.zr[(lx + m) * SP_RY + ly], acc);
    }
    L.zxr[k] = acc;
  }
  __syncthreads();
  int bad = 0;
  for (int k = threadIdx.x; k < FSMSP_TX * FSMSP_TY; k += blockDim.x) {
    int lx = k / FSMSP_TY, ly = k - lx * FSMSP_TY;
    int gx = x0 + lx, gy = y0 + ly;
    if (gx < nx && gy < ny) {
      const float *w = L.wy + ly * FSM_SPLINE_W;
      float zy = 0.0f, zxy = 0.0f;
#pragma unroll
      for (int m = 0; m < FSM_SPLINE_W; ++m) {
        zy = fmaf(w[m], L.zr[(lx + B) * SP_RY + ly + m], zy);
        zxy = fmaf(w[m], L.zxr[lx * SP_RY + ly + m], zxy);
      }
      float z = L.zr[(lx + B) * SP_RY + ly + B];
      float zx = L.zxr[lx * SP_RY + ly + B];
      ev.planes[(long)gx * ny + gy] = make_float4(z, zy, zx, zxy);
      if (ev.planes_h) {
        const float sc = L.sc;
        const float tol = -FSM_MFMA_POS_EPS * fmaxf(fmaxf(L.red[0], L.red[1]), fmaxf(L.red[2], L.red[3]));
        const float big = fmaxf(fmaxf(fabsf(z), fabsf(zy)), fmaxf(fabsf(zx), fabsf(zxy))) * sc;
        bad |= !(knot_cp_min(z, zy, zx, zxy) >= tol) || !(big <= FSM_MFMA_MAXABS);
        unsigned short h[4], l[4];
        split_half(z * sc, h[0], l[0]);
        split_half(zy * sc, h[1], l[1]);
        split_half(zx * sc, h[2], l[2]);
        split_half(zxy * sc, h[3], l[3]);
        ev.planes_h[(long)gx * ny + gy] =
            make_uint4(h[0] | ((unsigned)h[1] << 16), h[2] | ((unsigned)h[3] << 16), l[0] | ((unsigned)l[1] << 16),
                       l[2] | ((unsigned)l[3] << 16));
      }
    }
  }
  if (ev.planes_h) {
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) ev.sp_flags[tx * ev.sp_tiles_y + ty] = bad ? 1u : 0u;
  }
}

// Batched spline tile: FSMSP2_TX rows x FSMSP_TY columns of knots, with the
// same outputs as spline2d_tile bit for bit (each slope is the same
// sequential fmaf chain over the same band weights, and each FSMSP_TX x
// FSMSP_TY flag tile keeps its own positivity tolerance) but with the band
// products from registers: for z_x a lane owns one column and 16 rows and
// slides a 48-value window down the column, read from the knot values the
// knot pass wrote (zk); for z_y and z_xy a lane owns one row and 8 columns and
// slides a 40-value window along the row, from LDS.  A wave's rows (z_x) or
// columns (z_y) are uniform, so the weights are scalar loads, one set for the
// whole wave where its rows or columns all carry the interior weights
// ([cxlo, cxhi], [cylo, cyhi]).  Only the tile's own rows are staged in LDS
// (33.8 KiB per workgroup, four per CU).
#define SP2_RY (FSMSP_TY + 2 * FSM_SPLINE_BAND)
#define SP2_P (SP2_RY + 1)  // odd pitch: lanes walking rows do not share banks
#define SP2_OP (FSMSP_TY + 1)  // staged knots per tile row (float4): rows start in different banks
#define SP2_ZBUF (FSMSP2_TX * SP2_OP * 4 > 2 * FSMSP2_TX * SP2_P ? FSMSP2_TX * SP2_OP * 4 : 2 * FSMSP2_TX * SP2_P)
// The band weights are read-only for the kernel's lifetime: read through the
// constant address space, so uniform reads become scalar loads.
typedef __attribute__((address_space(4))) const float fsm_const_f;

struct __align__(16) SplineLds2 {
  // z over the tile rows and the y halo (zr), then z_x over the same (zxr);
  // afterwards the staged knots of the tile
  float zbuf[SP2_ZBUF];
  __device__ float *zr() { return zbuf; }
  __device__ float *zxr() { return zbuf + FSMSP2_TX * SP2_P; }
  float zmax[FSMSP2_TX / FSMSP_TX];  // largest |z| over each flag tile's band region
  unsigned bad[FSMSP2_TX / FSMSP_TX];
  float sc;
};

__device__ void spline2d_tile2(const Event2D &ev, int tx2, int ty, SplineLds2 &L) {
  constexpr int B = FSM_SPLINE_BAND, W = FSM_SPLINE_W, NF = FSMSP2_TX / FSMSP_TX;
  constexpr int RA = FSMSP2_TX / 4;       // z_x rows per wave
  constexpr int CB = FSMSP_TY / 4;        // z_y columns per wave
  static_assert(SP2_RY == 64 && FSMSP2_TX == 64, "one lane per staged column and per tile row");
  static_assert(RA == FSMSP_TX, "a wave's z_x rows are one flag tile, its window that tile's band region");
  const int nx = ev.nx, ny = ev.ny;
  const int x0 = tx2 * FSMSP2_TX, y0 = ty * FSMSP_TY;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // z_x (grid units) for tile rows x0 + RA wave + r and staged column lane
  // (y = y0 - B + lane), from the column's z over rows x0 + RA wave - B ...
  // x0 + RA wave + RA + B - 1
  {
    const int gy = y0 - B + lane;
    const bool yin = gy >= 0 && gy < ny;
    const int r0 = x0 + RA * wave;
    float win[RA + 2 * B];
#pragma unroll
    for (int u = 0; u < RA + 2 * B; ++u) {
      const int gx = r0 - B + u;
      win[u] = (yin && gx >= 0 && gx < nx) ? fsm_gload(ev.zk + (long)gx * ny + gy) : 0.0f;
    }
    __syncthreads();  // the previous item's reads of L are done
    if (ev.planes_h) {
      float m = 0.0f;
#pragma unroll
      for (int u = 0; u < RA + 2 * B; ++u) m = fmaxf(m, fabsf(win[u]));
      for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
      if (lane == 0) L.zmax[wave] = m;
      if (threadIdx.x < 64) {
        const float sc = knot_scale_from_parts(ev);
        if (threadIdx.x == 0) {
          L.sc = sc;
          ev.mscale[FSM_KS_PARTS] = sc;  // the same value from every block
        }
      }
      if (threadIdx.x < NF) L.bad[threadIdx.x] = 0u;
    }
#pragma unroll
    for (int r = 0; r < RA; ++r) L.zr()[(RA * wave + r) * SP2_P + lane] = win[r + B];
    if (r0 >= ev.cxlo && r0 + RA - 1 <= ev.cxhi) {
      fsm_const_f *w = (fsm_const_f *)ev.gx + (long)ev.cxlo * W;  // every row of this wave has these weights
#pragma unroll
      for (int r = 0; r < RA; ++r) {
        float acc = 0.0f;
#pragma unroll
        for (int m = 0; m < W; ++m) acc = fmaf(w[m], win[r + m], acc);
        L.zxr()[(RA * wave + r) * SP2_P + lane] = (yin && r0 + r < nx) ? acc : 0.0f;
      }
    } else {
#pragma unroll
      for (int r = 0; r < RA; ++r) {
        float acc = 0.0f;
        if (r0 + r < nx) {
          fsm_const_f *w = (fsm_const_f *)ev.gx + (long)(r0 + r) * W;
#pragma unroll
          for (int m = 0; m < W; ++m) acc = fmaf(w[m], win[r + m], acc);
        }
        L.zxr()[(RA * wave + r) * SP2_P + lane] = (yin && r0 + r < nx) ? acc : 0.0f;
      }
    }
  }
  __syncthreads();
  // z_y, z_xy for tile row lane (x = x0 + lane) and columns y0 + CB wave + j,
  // staged as knots {z, z_y, z_x, z_xy} in the LDS the windows came from
  float4 *stg = reinterpret_cast<float4 *>(L.zbuf);  // [FSMSP2_TX rows][SP2_OP]
  {
    float zw[CB + 2 * B], zxw[CB + 2 * B];
#pragma unroll
    for (int u = 0; u < CB + 2 * B; ++u) {
      zw[u] = L.zr()[lane * SP2_P + CB * wave + u];
      zxw[u] = L.zxr()[lane * SP2_P + CB * wave + u];
    }
    const int c0 = y0 + CB * wave;
    const bool canon = c0 >= ev.cylo && c0 + CB - 1 <= ev.cyhi;
    __syncthreads();  // every window is in registers: the staging may overwrite zr
#pragma unroll
    for (int j = 0; j < CB; ++j) {
      const int gy = c0 + j;
      if (gy >= ny) break;
      fsm_const_f *w = (fsm_const_f *)ev.gy + (long)(canon ? ev.cylo : gy) * W;
      float zy = 0.0f, zxy = 0.0f;
#pragma unroll
      for (int m = 0; m < W; ++m) {
        zy = fmaf(w[m], zw[j + m], zy);
        zxy = fmaf(w[m], zxw[j + m], zxy);
      }
      stg[lane * SP2_OP + CB * wave + j] = make_float4(zw[j + B], zy, zxw[j + B], zxy);
    }
  }
  __syncthreads();
  // Coalesced writes: element k of the 64 x 32 tile is row k / 32, column
  // k % 32, so each wave stores two whole 32-knot row runs per instruction.
  // The fp16 halves and the MFMA-safety test are made here from the staged
  // knot; a wave's two rows lie in one flag tile.
  {
    const float sc = ev.planes_h ? L.sc : 0.0f;
#pragma unroll
    for (int q = 0; q < FSMSP2_TX * FSMSP_TY / 256; ++q) {
      const int k = (int)threadIdx.x + 256 * q;
      const int row = k / FSMSP_TY, col = k % FSMSP_TY;
      const int gx = x0 + row, gy = y0 + col;
      const bool live = gx < nx && gy < ny;
      bool bad = false;
      if (live) {
        const float4 v = stg[row * SP2_OP + col];
        fsm_gstore(ev.planes + (long)gx * ny + gy, v);
        if (ev.planes_h) {
          const float tol = -FSM_MFMA_POS_EPS * L.zmax[row / FSMSP_TX];
          const float big = fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))) * sc;
          bad = !(knot_cp_min(v.x, v.y, v.z, v.w) >= tol) || !(big <= FSM_MFMA_MAXABS);
          unsigned short h[4], l[4];
          split_half(v.x * sc, h[0], l[0]);
          split_half(v.y * sc, h[1], l[1]);
          split_half(v.z * sc, h[2], l[2]);
          split_half(v.w * sc, h[3], l[3]);
          fsm_gstore(ev.planes_h + (long)gx * ny + gy,
                     make_uint4(h[0] | ((unsigned)h[1] << 16), h[2] | ((unsigned)h[3] << 16),
                                l[0] | ((unsigned)l[1] << 16), l[2] | ((unsigned)l[3] << 16)));
        }
      }
      if (ev.planes_h && __ballot(bad) != 0ull && lane == 0) atomicOr(&L.bad[row / FSMSP_TX], 1u);
    }
  }
  if (ev.planes_h) {
    __syncthreads();
    const int sx = tx2 * NF + (int)threadIdx.x;
    if ((int)threadIdx.x < NF && sx * FSMSP_TX < nx) ev.sp_flags[sx * ev.sp_tiles_y + ty] = L.bad[threadIdx.x];
  }
}

// Knot scale of the MFMA form: 2^k <= FSM_MFMA_ZTARGET / max|z| over the
// event (z as the spline kernel computes it).  Pass 1 (k_knot_scale): part p
// of FSM_KS_PARTS writes the largest |z| of its contiguous chunk to
// mscale[p]; pass 2 (the spline kernel's blocks) reduce the parts and each
// writes the same scale to mscale[FSM_KS_PARTS].  No pre-cleared memory and
// no atomics, so both passes may be captured in a graph.
__device__ void knot_scale_part(const Event2D &ev, int part, float *red) {
  const long n = (long)ev.nx * ev.ny;
  const long k0 = n * part / FSM_KS_PARTS, k1 = n * (part + 1) / FSM_KS_PARTS;
  float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f, m3 = 0.0f;
  const long step = blockDim.x;
  long k = k0 + threadIdx.x;
  float *zk = ev.zk;  // the knot values, kept for the batch spline kernel
  for (; k + 3 * step < k1; k += 4 * step) {
    float z[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) z[j] = knot_z(ev, fsm_gload(ev.e_gev + k + j * step));
    if (zk)
#pragma unroll
      for (int j = 0; j < 4; ++j) fsm_gstore_at(zk, (unsigned)(k + j * step), z[j]);
    m0 = fmaxf(m0, fabsf(z[0]));
    m1 = fmaxf(m1, fabsf(z[1]));
    m2 = fmaxf(m2, fabsf(z[2]));
    m3 = fmaxf(m3, fabsf(z[3]));
  }
  for (; k < k1; k += step) {
    const float z = knot_z(ev, fsm_gload(ev.e_gev + k));
    if (zk) fsm_gstore_at(zk, (unsigned)k, z);
    m0 = fmaxf(m0, fabsf(z));
  }
  float m = fmaxf(fmaxf(m0, m1), fmaxf(m2, m3));
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmaxf(m, red[w]);
    ev.mscale[part] = m;
  }
}

// Scale from the parts (every lane of a wave returns it).
__device__ float knot_scale_from_parts(const Event2D &ev) {
  float m = 0.0f;
  for (int p = threadIdx.x & 63; p < FSM_KS_PARTS; p += 64) m = fmaxf(m, ev.mscale[p]);
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  float s = 1.0f;
  if (m > 0.0f && m < INFINITY) {
    int ex = 0;
    frexpf(FSM_MFMA_ZTARGET / m, &ex);
    s = ldexpf(1.0f, min(max(ex - 1, -60), 60));
  }
  return s;
}

__global__ __launch_bounds__(256) void k_knot_scale_one(Event2D ev) {
  __shared__ float red[4];
  knot_scale_part(ev, blockIdx.x, red);
}

// Batched: item = (event, part), for every 2-D event of a batch (the knot
// values go to zk for the batch spline kernel; the maxima are used on the
// MFMA path only).
__global__ __launch_bounds__(256) void k_knot_scale_batch(const Event2D *__restrict__ evs,
                                                          const int4 *__restrict__ items, int n_items) {
  __shared__ float red[4];
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const int4 item = fsm_sload(items + xcd_order(it, n_items));
    const Event2D ev = fsm_sload(evs + item.x);
    knot_scale_part(ev, item.y, red);
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_spline2d_one(Event2D ev) {
  __shared__ SplineLds L;
  int tiles_y = (ev.ny + FSMSP_TY - 1) / FSMSP_TY;
  const int t = xcd_order(blockIdx.x, gridDim.x);
  spline2d_tile(ev, t / tiles_y, t % tiles_y, L);
}

// Batched: item = (event, FSMSP2_TX-row tile, FSMSP_TY-column tile).
__global__ __launch_bounds__(256) void k_spline2d_batch(const Event2D *__restrict__ evs,
                                                        const int4 *__restrict__ items, int n_items) {
  __shared__ SplineLds2 L;
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    // descriptors through scalar loads (uniform; written before the launch)
    const int4 item = fsm_sload(items + xcd_order(it, n_items));
    const Event2D ev = fsm_sload(evs + item.x);
    spline2d_tile2(ev, item.y, item.z, L);
  }
}

// The 2-D streaming + quadrature kernels live in fsm_stream2d.hip.

// ======================================================================
// 3-D source planes and streaming
// ======================================================================
// e_gev (reference layout [nxy][neta]) -> src (eta-major [neta][nxy]) with
// src = e0 * norm_factor (src/FreeStream.cpp:346, first product).
__device__ void src3d_tile(const float *__restrict__ e_gev, float *__restrict__ src, int nxy, int neta,
                           float norm_factor, int raw, int bx, int by, float (*tile)[65]) {
  const int xy0 = bx * 64, et0 = by * 64;
  const int tx = threadIdx.x & 63, tyb = threadIdx.x >> 6;
  __syncthreads();
  for (int r = tyb; r < 64; r += 4) {
    int xy = xy0 + r, et = et0 + tx;
    float v = 0.0f;
    if (xy < nxy && et < neta) {
      const float x = e_gev[(long)xy * neta + et];
      // src/FreeStream.cpp:384 (n_B * norm) or :346 (e0 * norm); the
      // rapidity factor is applied per slot.  `raw`: the input is already
      // the density (n_B, or an fm^-4 energy profile)
      v = (raw ? x : ingest_e0(x)) * norm_factor;
    }
    tile[r][tx] = v;
  }
  __syncthreads();
  for (int r = tyb; r < 64; r += 4) {
    int et = et0 + r, xy = xy0 + tx;
    if (xy < nxy && et < neta) src[(long)et * nxy + xy] = tile[tx][r];
  }
}

__global__ __launch_bounds__(256) void k_src3d(const float *__restrict__ e_gev, float *__restrict__ src,
                                               int nxy, int neta, float norm_factor, int raw) {
  __shared__ float tile[64][65];
  src3d_tile(e_gev, src, nxy, neta, norm_factor, raw, blockIdx.x, blockIdx.y, tile);
}

// Batched: item = (event, xy tile, eta tile).
__global__ __launch_bounds__(256) void k_src3d_batch(const Event3D *__restrict__ evs,
                                                     const int4 *__restrict__ items, int n_items) {
  __shared__ float tile[64][65];
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const int4 item = fsm_sload(items + xcd_order(it, n_items));
    const Event3D ev = fsm_sload(evs + item.x);
    src3d_tile(ev.e_gev, ev.src, ev.nx * ev.ny, ev.neta, ev.norm_factor, ev.charge | ev.raw_e, item.y, item.z,
               tile);
  }
}

// Generic plane transpose: in [planes][rows][cols] -> out [planes][cols][rows].
__global__ __launch_bounds__(256) void k_transpose(const float *__restrict__ in, float *__restrict__ out,
                                                   int rows, int cols, long in_plane, long out_plane) {
  __shared__ float tile[64][65];
  const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64, p = blockIdx.z;
  const int tx = threadIdx.x & 63, tyb = threadIdx.x >> 6;
  const float *ip = in + p * in_plane;
  float *op = out + p * out_plane;
  for (int r = tyb; r < 64; r += 4) {
    int rr = r0 + r, cc = c0 + tx;
    tile[r][tx] = (rr < rows && cc < cols) ? ip[(long)rr * cols + cc] : 0.0f;
  }
  __syncthreads();
  for (int r = tyb; r < 64; r += 4) {
    int cc = c0 + r, rr = r0 + tx;
    if (rr < rows && cc < cols) op[(long)cc * rows + rr] = tile[tx][r];
  }
}

// One transverse tile (FSM3D_TILE consecutive cells of the ix*NY+iy order,
// i.e. a few whole rows) of one eta plane.  Per rapidity slot the kernel walks
// the slot's live azimuth samples (host tables: offset, bilinear weights and
// the box of cells the reference keeps).  The source the samples read is the
// eta-blended plane
//   Q(x, y) = (1-f) rf0 E[eta0](x, y) + f rf1 E[eta0+1](x, y),
// staged in LDS as a window of whole rows (pitch NY, rows outside the grid
// zero).  The host plans the windows of every (slot, tile) pair: a window
// holds the source rows that a run of consecutive samples reads for this
// tile, so a slot whose samples only graze the grid stages only what they
// read.  Every read lies in the window: a kept cell's source rows ix+ox,
// ix+ox+1 lie in the window's rows and its columns iy+oy, iy+oy+1 in
// [1, NY-1] (the kept box; checked on the host).  With pitch NY consecutive
// cells of the tile read consecutive LDS words across a row boundary too.
//
// Stages window rows [lo, lo + nr) of Q = a0 E0 + a1 E1 (a0 = (1-f) rf0,
// a1 = f rf1).  Lane l of wave w covers columns l, l + 64, ... of rows w,
// w + 4, ...; four rows' loads are issued before their first use.
__device__ __forceinline__ void stage_q_rows(float *q, const float *__restrict__ p0, const float *__restrict__ p1,
                                             float a0, float a1, int lo, int nr, int nx, int ny) {
  constexpr int RU = 4;
  const int lane = (int)threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  // byte offsets from a wave-uniform base: 32-bit address arithmetic
  const FSM_AS_GLOBAL char *g0 = (const FSM_AS_GLOBAL char *)p0;
  const FSM_AS_GLOBAL char *g1 = (const FSM_AS_GLOBAL char *)p1;
  for (int c = lane; c - lane < ny; c += 64) {
    const bool cin = c < ny;
    for (int r0 = wave; r0 < nr; r0 += 4 * RU) {
      float e0[RU], e1[RU];
      bool ok[RU];
#pragma unroll
      for (int u = 0; u < RU; ++u) {
        const int gx = lo + r0 + 4 * u;  // wave-uniform
        ok[u] = cin && r0 + 4 * u < nr && (unsigned)gx < (unsigned)nx;
        // an element off the grid loads element 0 and stores 0
        const unsigned ob = ok[u] ? 4u * (unsigned)(gx * ny + c) : 0u;
        e0[u] = *(const FSM_AS_GLOBAL float *)(g0 + ob);
        e1[u] = *(const FSM_AS_GLOBAL float *)(g1 + ob);
      }
#pragma unroll
      for (int u = 0; u < RU; ++u) {
        const int r = r0 + 4 * u;
        if (cin && r < nr) q[r * ny + c] = ok[u] ? fmaf(a0, e0[u], a1 * e1[u]) : 0.0f;
      }
    }
  }
  // row nr is zero: a kept cell whose rounding-decided floor index sits one
  // column left of the offset's may read column NY with weight 0, i.e. the
  // next row's column 0 (a staged row, or this one)
  for (int c = (int)threadIdx.x; c < ny; c += FSM3D_THREADS) q[nr * ny + c] = 0.0f;
}

typedef unsigned short fsm_u16x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) const float fsm_lds_cf;  // 32-bit LDS addresses

// Bilinear values of one sample at the lane's FSM3D_M cells: a cell outside
// the sample's kept box (one packed 16-bit subtract/min/compare of
// (ix << 16 | iy) against the box) reads the 2 x 2 zero patch after the
// window, so its value needs no select.
struct Q3Taps {
  float q00[FSM3D_M], q01[FSM3D_M], q10[FSM3D_M], q11[FSM3D_M];
};
__device__ __forceinline__ void q3_taps(const Samp3D &P, unsigned wb, const unsigned (&pk)[FSM3D_M],
                                        const unsigned (&ba)[FSM3D_M], unsigned zero_b, unsigned pitch_b,
                                        Q3Taps &q) {
  const unsigned delta_b = (unsigned)P.d0 - wb;  // modulo 2^32
#pragma unroll
  for (int m = 0; m < FSM3D_M; ++m) {
    const fsm_u16x2 t = __builtin_bit_cast(fsm_u16x2, pk[m]) - __builtin_bit_cast(fsm_u16x2, P.mlo);
    const fsm_u16x2 mn = __builtin_elementwise_min(t, __builtin_bit_cast(fsm_u16x2, P.mext));
    const bool ok = __builtin_bit_cast(unsigned, mn) == __builtin_bit_cast(unsigned, t);
    const unsigned ad = ok ? ba[m] + delta_b : zero_b;
    fsm_lds_cf *qa = (fsm_lds_cf *)(size_t)ad;
    fsm_lds_cf *qc = (fsm_lds_cf *)(size_t)(ad + pitch_b);
    q.q00[m] = qa[0];
    q.q01[m] = qa[1];
    q.q10[m] = qc[0];
    q.q11[m] = qc[1];
  }
}
// Adds one sample's value v at the lane's cells to the angular moments
// v, v cos, v sin, v cos^2, v cos sin (v sin^2 = v - v cos^2 at the end).
__device__ __forceinline__ void q3_moments(const Samp3D &P, const Q3Taps &q, float (&mo)[FSM3D_M][5]) {
#pragma unroll
  for (int m = 0; m < FSM3D_M; ++m) {
    float v = q.q00[m] * P.w00;
    v = fmaf(q.q01[m], P.w01, v);
    v = fmaf(q.q10[m], P.w10, v);
    v = fmaf(q.q11[m], P.w11, v);
    mo[m][0] += v;
    mo[m][1] = fmaf(v, P.c, mo[m][1]);
    mo[m][2] = fmaf(v, P.s, mo[m][2]);
    mo[m][3] = fmaf(v, P.cc, mo[m][3]);
    mo[m][4] = fmaf(v, P.cs, mo[m][4]);
  }
}

__device__ void stream3d_tile(const Event3D &ev, int tile, int ieta, float *q) {
  const int nx = ev.nx, ny = ev.ny, nxy = nx * ny;
  const int pitch = ny;
  const int W = ev.win_rows;
  const int c0 = tile * FSM3D_TILE;
  // the lane's cells c0 + tid + 256 m as (ix << 16 | iy) (all ones past the
  // grid: fails every kept-box test) and as LDS byte address in the window at
  // zero row offset; a dropped cell reads the 2 x 2 zero patch after the
  // window (rows W, W+1, columns 0, 1)
  const unsigned qbase = (unsigned)(size_t)(fsm_lds_cf *)q;
  unsigned pk[FSM3D_M];
  unsigned ba[FSM3D_M];
#pragma unroll
  for (int m = 0; m < FSM3D_M; ++m) {
    const int c = c0 + threadIdx.x + FSM3D_THREADS * m;
    const int cx = c / ny, cy = c - cx * ny;
    pk[m] = c < nxy ? ((unsigned)cx << 16) | (unsigned)cy : 0xFFFFFFFFu;
    ba[m] = qbase + (c < nxy ? 4u * (unsigned)(cx * pitch + cy) : 0u);
  }
  const unsigned zero_b = qbase + 4u * (unsigned)(W * pitch);
  const unsigned pitch_b = 4u * (unsigned)pitch;
  float T[FSM3D_M][10];
#pragma unroll
  for (int m = 0; m < FSM3D_M; ++m)
#pragma unroll
    for (int v = 0; v < 10; ++v) T[m][v] = 0.0f;

  for (int k = 0; k < ev.nrap; ++k) {
    const RapSlot S = fsm_sload(ev.slots + (long)ieta * ev.nrap + k);
    if (S.ns == 0) continue;
    const int2 wr = fsm_sload(ev.srange + ((long)ieta * ev.nrap + k) * ev.tiles + tile);
    if (wr.x >= wr.y) continue;  // no sample keeps a cell of this tile
    const float a0 = (1.0f - S.feta) * S.rf0, a1 = S.feta * S.rf1;
    const float *p0 = ev.src + (long)S.eta0 * nxy;
    const float *p1 = ev.src + (long)(S.eta0 + 1) * nxy;
    float mo[FSM3D_M][5];
#pragma unroll
    for (int m = 0; m < FSM3D_M; ++m)
#pragma unroll
      for (int c = 0; c < 5; ++c) mo[m][c] = 0.0f;
    const Samp3D *sp = ev.samp + S.s0;
    for (int w = wr.x; w < wr.y; ++w) {
      const int4 win = fsm_sload(ev.win + w);  // samples [x, y), source rows [z, z + w)
      __syncthreads();  // previous window no longer read
      stage_q_rows(q, p0, p1, a0, a1, win.z, win.w, nx, ny);
      if (threadIdx.x < 4) q[W * pitch + (threadIdx.x & 1) + (threadIdx.x >> 1) * pitch] = 0.0f;
      __syncthreads();
      const unsigned wb = 4u * (unsigned)(win.z * pitch);
      // two samples per step: the second one's LDS reads are in flight while
      // the first one's moments are summed (the order of the sums is kept)
      int j = win.x;
      for (; j + 1 < win.y; j += 2) {
        const Samp3D P0 = fsm_sload(sp + j);
        const Samp3D P1 = fsm_sload(sp + j + 1);
        Q3Taps qa, qb;
        q3_taps(P0, wb, pk, ba, zero_b, pitch_b, qa);
        q3_taps(P1, wb, pk, ba, zero_b, pitch_b, qb);
        q3_moments(P0, qa, mo);
        q3_moments(P1, qb, mo);
      }
      if (j < win.y) {
        const Samp3D P0 = fsm_sload(sp + j);
        Q3Taps qa;
        q3_taps(P0, wb, pk, ba, zero_b, pitch_b, qa);
        q3_moments(P0, qa, mo);
      }
    }
    // hypertrig rows (src/LandauMatch.cpp:46-55) times the Jacobian (:93,101)
#pragma unroll
    for (int m = 0; m < FSM3D_M; ++m) {
      const float j0 = S.jac;
      T[m][0] = fmaf(j0, mo[m][0], T[m][0]);
      T[m][1] = fmaf(j0 * S.inv_ch, mo[m][1], T[m][1]);
      T[m][2] = fmaf(j0 * S.inv_ch, mo[m][2], T[m][2]);
      T[m][3] = fmaf(j0 * S.th_t, mo[m][0], T[m][3]);
      T[m][4] = fmaf(j0 * S.inv_ch2, mo[m][3], T[m][4]);
      T[m][5] = fmaf(j0 * S.inv_ch2, mo[m][4], T[m][5]);
      T[m][6] = fmaf(j0 * S.th_tch, mo[m][1], T[m][6]);
      T[m][7] = fmaf(j0 * S.inv_ch2, mo[m][0] - mo[m][3], T[m][7]);
      T[m][8] = fmaf(j0 * S.th_tch, mo[m][2], T[m][8]);
      T[m][9] = fmaf(j0 * S.th2_t2, mo[m][0], T[m][9]);
    }
  }
  const long P = ev.t_plane;
#pragma unroll
  for (int m = 0; m < FSM3D_M; ++m) {
    const int c = c0 + threadIdx.x + FSM3D_THREADS * m;
    if (c >= nxy) continue;
    const long idx = (long)ieta * nxy + c;
#pragma unroll
    for (int v = 0; v < 10; ++v) fsm_gstore(ev.T + v * P + idx, T[m][v] * ev.d_phip);  // :105
  }
}

__global__ __launch_bounds__(FSM3D_THREADS, 4) void k_stream3d_one(Event3D ev) {
  extern __shared__ float qlds[];
  // eta planes in decreasing order of work, so the last workgroups to start
  // are the short ones; tiles of one plane round-robin over the XCDs
  stream3d_tile(ev, blockIdx.x % ev.tiles, ev.eta_order[blockIdx.x / ev.tiles], qlds);
}

__global__ __launch_bounds__(FSM3D_THREADS, 4) void k_stream3d_batch(const Event3D *__restrict__ evs,
                                                                     const int4 *__restrict__ items,
                                                                     int n_items) {
  extern __shared__ float qlds[];
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    // descriptors through scalar loads: pointers read from memory are generic
    // and their loads would otherwise go through flat vector loads
    const int4 item = fsm_sload(items + it);  // the host orders items by decreasing work
    const Event3D ev = fsm_sload(evs + item.x);
    stream3d_tile(ev, item.y, item.z, qlds);
  }
}

// ======================================================================
// Landau matching + EoS + viscous parts + output pack: the per-cell code is
// in fsm_landau.h (shared with the 2-D stream kernels' fused epilogue)
// ======================================================================
#include "fsm_landau.h"

// One cell per thread, reference index is (T read through the layout map);
// workgroup-uniform call.
template <int N>
__device__ void landau_cell_out(const EventLandau &ev, bool live, long is) {
  // cell indices fit 32 bits (DIM is an int): 32-bit divisions
  long tidx = is;
  if (ev.t_eta_major) {
    const unsigned xy = (unsigned)is / (unsigned)ev.neta;
    const unsigned ie = (unsigned)is - xy * (unsigned)ev.neta;
    tidx = (long)ie * ev.nxy + xy;
  } else if (ev.t_row) {
    const unsigned ix = (unsigned)is / (unsigned)ev.t_ny;
    tidx = (long)ix * ev.t_row + ((unsigned)is - ix * (unsigned)ev.t_ny);
  }
  CellOut r;
  landau_cell_compute<N>(ev, live, tidx, r);
  if (!live) return;
  if (ev.out) {
#pragma unroll
    for (int c = 0; c < FSM_NUM_OUTPUTS; ++c)
      fsm_gstore_at(ev.out + c * ev.out_plane, (unsigned)is, plane_double(ev, cell_value(r, c), c));
  }
  if (ev.J) {
#pragma unroll
    for (int c = FSM_NUM_OUTPUTS; c < FSM_NUM_OUTPUTS + 5; ++c)
      fsm_gstore_at(ev.out_b + (c - FSM_NUM_OUTPUTS) * ev.out_plane, (unsigned)is, plane_double(ev, cell_value(r, c), c));
  }
  cell_store_float(ev, is, r);
}

// 3-D events (T eta-major): a tile of FSML_TE eta planes x FSML_TX transverse
// cells.  T is read along the transverse index (coalesced in the eta-major
// layout); the outputs go through LDS and leave along eta, the fast index of
// the reference layout, so both sides move whole cache-line runs.
__device__ void landau3d_tile(const EventLandau &ev, int xt, int et, float *lds) {
  const int xl = threadIdx.x % FSML_TX, el = threadIdx.x / FSML_TX;
  const long xy = (long)xt * FSML_TX + xl;
  const int ie = et * FSML_TE + el;
  const bool live = xy < ev.nxy && ie < ev.neta;
  const int np = ev.J ? FSM_NUM_OUTPUTS + 5 : FSM_NUM_OUTPUTS;
  CellOut r;
  // The trailing barrier below keeps the staged outputs of one tile from
  // being overwritten by the next tile of the same workgroup.
  landau_cell_compute<4>(ev, live, (long)ie * ev.nxy + xy, r);
  if (live) {
    const int li = xl * FSML_PITCH + el;
#pragma unroll
    for (int c = 0; c < FSM_NUM_OUTPUTS + 5; ++c)
      if (c < np) lds[c * FSML_PLANE + li] = cell_value(r, c);
    cell_store_float(ev, xy * ev.neta + ie, r);
  }
  __syncthreads();
  const int wxl = threadIdx.x / FSML_TE, wel = threadIdx.x % FSML_TE;
  const long wxy = (long)xt * FSML_TX + wxl;
  const int wie = et * FSML_TE + wel;
  if (wxy < ev.nxy && wie < ev.neta) {
    const long is = wxy * ev.neta + wie;
    const int li = wxl * FSML_PITCH + wel;
    for (int c = (ev.out ? 0 : FSM_NUM_OUTPUTS); c < np; ++c) {
      double *o = c < FSM_NUM_OUTPUTS ? ev.out + c * ev.out_plane : ev.out_b + (c - FSM_NUM_OUTPUTS) * ev.out_plane;
      fsm_gstore_at(o, (unsigned)is, plane_double(ev, lds[c * FSML_PLANE + li], c));
    }
  }
  __syncthreads();  // every lane has read its staged outputs
}

// One stage of the reference's per-cell chain on caller-given inputs
// (calculatePressure, calculateBulkPressure or calculateShearViscTensor):
// one cell per thread over planes of `dim` floats.
__global__ __launch_bounds__(256) void k_stage_cells(StageCells a) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long is = (long)blockIdx.x * blockDim.x + threadIdx.x; is < a.dim; is += stride) {
    if (a.stage == FSM_STAGE_PRESSURE) {
      a.p_out[is] = eos_pressure(a.e[is], a.eos_type);
      continue;
    }
    float t[10];
#pragma unroll
    for (int c = 0; c < 10; ++c) t[c] = a.T[c * a.dim + is];
    if (a.stage == FSM_STAGE_BULK) {
      a.Pi_out[is] = bulk_pressure(t, a.tau, a.e[is], a.p[is]);
      continue;
    }
    float u[4], pi[10];
#pragma unroll
    for (int c = 0; c < 4; ++c) u[c] = a.u[c * a.dim + is];
    shear_tensor(t, a.inv_tau2, a.e[is], u, a.p[is], a.Pi[is], pi);
#pragma unroll
    for (int c = 0; c < 10; ++c) a.pi_out[c * a.dim + is] = pi[c];
  }
}

template <int N>
__global__ __launch_bounds__(256) void k_landau_one(EventLandau ev) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long base = (long)blockIdx.x * blockDim.x; base < ev.dim; base += stride) {
    const long is = base + threadIdx.x;
    landau_cell_out<N>(ev, is < ev.dim, is);
  }
}

__global__ __launch_bounds__(256) void k_landau3d_one(EventLandau ev) {
  extern __shared__ float lds_out[];
  // eta tiles fastest: the workgroups running together cover whole eta runs
  // of a few transverse cells, i.e. contiguous stretches of every output
  // plane (eta is the fast index of the reference layout)
  const int etiles = (ev.neta + FSML_TE - 1) / FSML_TE;
  const int b = xcd_order(blockIdx.x, gridDim.x);
  landau3d_tile(ev, b / etiles, b % etiles, lds_out);
}

// Batched 3-D events (eta-major T): item = (event, transverse tile, eta
// tile, 1).
__global__ __launch_bounds__(256) void k_landau_batch(const EventLandau *__restrict__ evs,
                                                      const int4 *__restrict__ items, int n_items) {
  extern __shared__ float lds_out[];
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const int4 item = fsm_sload(items + xcd_order(it, n_items));
    const EventLandau ev = fsm_sload(evs + item.x);
    landau3d_tile(ev, item.y, item.z, lds_out);
  }
}

// Batched 2-D events whose tensors come from the 2-D stream kernel (eta
// components exactly zero): item = (event, cell offset, cell count, 0), the
// 3x3 Landau frame.
__global__ __launch_bounds__(256) void k_landau2d_batch(const EventLandau *__restrict__ evs,
                                                        const int4 *__restrict__ items, int n_items) {
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const int4 item = fsm_sload(items + xcd_order(it, n_items));
    const EventLandau ev = fsm_sload(evs + item.x);
    for (int base = 0; base < item.z; base += blockDim.x)
      landau_cell_out<3>(ev, base + (int)threadIdx.x < item.z, (long)item.y + base + threadIdx.x);
  }
}

}  // namespace

// ======================================================================
// Launchers (host)
// ======================================================================
hipError_t fsm_launch_spline2d(const Event2D &ev, hipStream_t s) {
  if (ev.planes_h) {
    hipLaunchKernelGGL(k_knot_scale_one, dim3(FSM_KS_PARTS), dim3(256), 0, s, ev);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  int tiles = ((ev.nx + FSMSP_TX - 1) / FSMSP_TX) * ((ev.ny + FSMSP_TY - 1) / FSMSP_TY);
  hipLaunchKernelGGL(k_spline2d_one, dim3(tiles), dim3(256), 0, s, ev);
  return hipGetLastError();
}

hipError_t fsm_launch_src3d(const Event3D &ev, hipStream_t s) {
  int nxy = ev.nx * ev.ny;
  dim3 grid((nxy + 63) / 64, (ev.neta + 63) / 64);
  hipLaunchKernelGGL(k_src3d, grid, dim3(256), 0, s, ev.e_gev, ev.src, nxy, ev.neta, ev.norm_factor,
                     ev.charge | ev.raw_e);
  return hipGetLastError();
}

hipError_t fsm_launch_stream3d(const Event3D &ev, hipStream_t s) {
  size_t lds = (size_t)ev.lds_floats * sizeof(float);
  hipLaunchKernelGGL(k_stream3d_one, dim3(ev.tiles * ev.neta), dim3(FSM3D_THREADS), lds, s, ev);
  return hipGetLastError();
}

hipError_t fsm_launch_stage_cells(const StageCells &a, hipStream_t s) {
  if (a.dim <= 0) return hipSuccess;
  const int blocks = (int)std::min<long>((a.dim + 255) / 256, 16384L);
  hipLaunchKernelGGL(k_stage_cells, dim3(blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t fsm_launch_transpose(const float *in, float *out, int rows, int cols, int planes, long in_plane,
                                long out_plane, hipStream_t s) {
  dim3 grid((rows + 63) / 64, (cols + 63) / 64, planes);
  hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, s, in, out, rows, cols, in_plane, out_plane);
  return hipGetLastError();
}

hipError_t fsm_launch_landau(const EventLandau &ev, hipStream_t s) {
  if (ev.t_eta_major && ev.neta > 1) {
    const int xtiles = (ev.nxy + FSML_TX - 1) / FSML_TX, etiles = (ev.neta + FSML_TE - 1) / FSML_TE;
    hipLaunchKernelGGL(k_landau3d_one, dim3(xtiles * etiles), dim3(256), fsm_landau_lds_bytes(ev.J != nullptr), s,
                       ev);
    return hipGetLastError();
  }
  int blocks = (int)std::min<long>((ev.dim + 255) / 256, 65535L);
  if (ev.eta_zero)
    hipLaunchKernelGGL(k_landau_one<3>, dim3(blocks), dim3(256), 0, s, ev);
  else
    hipLaunchKernelGGL(k_landau_one<4>, dim3(blocks), dim3(256), 0, s, ev);
  return hipGetLastError();
}

hipError_t fsm_set_lds_limits(int bytes2d, int bytes3d) {
  hipError_t e = fsm_stream2d_set_lds_limit(bytes2d);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute((const void *)k_stream3d_one, hipFuncAttributeMaxDynamicSharedMemorySize, bytes3d);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void *)k_stream3d_batch, hipFuncAttributeMaxDynamicSharedMemorySize, bytes3d);
}

// ---- batched launchers: one launch per stage for every event of a batch ----
static inline int batch_grid(int n_items) { return std::max(1, std::min(n_items, 16384)); }

hipError_t fsm_launch_knot_scale_batch(const Event2D *evs, const int4 *items, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_knot_scale_batch, dim3(batch_grid(n)), dim3(256), 0, s, evs, items, n);
  return hipGetLastError();
}

hipError_t fsm_launch_spline2d_batch(const Event2D *evs, const int4 *items, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_spline2d_batch, dim3(batch_grid(n)), dim3(256), 0, s, evs, items, n);
  return hipGetLastError();
}

hipError_t fsm_launch_src3d_batch(const Event3D *evs, const int4 *items, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_src3d_batch, dim3(batch_grid(n)), dim3(256), 0, s, evs, items, n);
  return hipGetLastError();
}

hipError_t fsm_launch_stream3d_batch(const Event3D *evs, const int4 *items, int n, size_t lds_bytes, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_stream3d_batch, dim3(batch_grid(n)), dim3(FSM3D_THREADS), lds_bytes, s, evs, items, n);
  return hipGetLastError();
}

size_t fsm_landau_lds_bytes(int baryon) {
  return (size_t)(baryon ? FSM_NUM_OUTPUTS + 5 : FSM_NUM_OUTPUTS) * FSML_PLANE * sizeof(float);
}

hipError_t fsm_launch_landau_batch(const EventLandau *evs, const int4 *items, int n, size_t lds_bytes,
                                   hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_landau_batch, dim3(batch_grid(n)), dim3(256), lds_bytes, s, evs, items, n);
  return hipGetLastError();
}

hipError_t fsm_launch_landau2d_batch(const EventLandau *evs, const int4 *items, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_landau2d_batch, dim3(batch_grid(n)), dim3(256), 0, s, evs, items, n);
  return hipGetLastError();
}
