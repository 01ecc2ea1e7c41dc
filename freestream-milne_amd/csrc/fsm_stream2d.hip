// fsm_stream2d.hip -- 2-D free streaming fused with the momentum quadrature
// (FP32 VALU path).
//
// freeStream in 2-D (src/FreeStream.cpp:156-248) samples the initial density
// at x - DTAU (cos phi_j, sin phi_j) with a bicubic spline (bilinear where the
// spline is <= 0), and calculateStressTensor (src/LandauMatch.cpp:61-108)
// sums the samples against the hypertrig rows (src/LandauMatch.cpp:46-55),
// which in 2-D are 1, c, s, c^2, cs, s^2 (cosh(0) = 1, tanh(0) = 0).  The
// shift is the same vector for every cell, so per phi_j the knot offset and
// the Hermite/bilinear weights are constants (Stencil2D).
//
// Work decomposition.  A workgroup of 4 waves owns a tile of 32 (y) x TX (x)
// cells.  In each wave, lane l covers y = l % 32 and a run of M consecutive x
// cells in half-wave l / 32; the M+1 knot columns a run touches are read once
// from LDS and shared by its M samples (the y-direction Hermite combination
// is done per knot column).  JS waves split the azimuth samples of one cell
// run between them and reduce through LDS at the end (JS = 4 quadruples the
// workgroups of an event for single-event latency).  For each contiguous
// range of phi_j (a StencilGroup) the knot box the tile reaches is staged in
// LDS; tiles whose box lies inside the grid by two cells skip the edge masks.
#include <hip/hip_runtime.h>
#include <math.h>

#include "fsm_internal.h"
#include "fsm_kernels.h"

namespace {

__device__ __forceinline__ bool ref_mask_ok2(int i, float shift, float mn, float h, int n) {
  const float x = __fadd_rn(__fmul_rn((float)i, h), mn);
  const float xn = __fsub_rn(x, shift);
  const float f = floorf(__fdiv_rn(__fsub_rn(xn, mn), h));
  return (f >= 1.0f) && (f < (float)(n - 1));
}

// src/FreeStream.cpp:213 mask; exact float re-evaluation next to the boundary.
__device__ __forceinline__ bool mask_ok2(int i, int g, float shift, float mn, float h, int n) {
  if ((unsigned)g <= 1u || (unsigned)(g - (n - 2)) <= 1u) return ref_mask_ok2(i, shift, mn, h, n);
  return (unsigned)(g - 1) <= (unsigned)(n - 3);
}

template <int M, int JS>
__device__ void stream2d_tile_v2(const Event2D &ev, int tile_x, int tile_y, float4 *lds) {
  constexpr int CG = 4 / JS;          // cell groups per workgroup
  constexpr int TX = CG * 2 * M;      // tile width in x
  constexpr int TY = 32;
  const int nx = ev.nx, ny = ev.ny;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int js = wave % JS, cg = wave / JS;
  const int yl = lane & 31, xh = lane >> 5;
  const int x0 = tile_x * TX, y0 = tile_y * TY;
  const int iy = y0 + yl;
  const int xr = (cg * 2 + xh) * M;  // first x of this lane's run, relative to x0
  const int ixb = x0 + xr;
  float acc[M][5];
#pragma unroll
  for (int m = 0; m < M; ++m)
#pragma unroll
    for (int c = 0; c < 5; ++c) acc[m][c] = 0.0f;

  const int xlast = min(x0 + TX - 1, nx - 1), ylast = min(y0 + TY - 1, ny - 1);
  for (int g = 0; g < ev.ngroups; ++g) {
    const StencilGroup G = ev.groups[g];
    const int LX = TX + G.bx1 - G.bx0;
    const int LY = TY + G.by1 - G.by0;
    // every sample of every real cell of the tile lands in [2, n-3]
    const bool interior = (x0 + G.bx0 >= 2) && (xlast + G.bx1 - 1 <= nx - 3) && (y0 + G.by0 >= 2) &&
                          (ylast + G.by1 - 1 <= ny - 3);
    const int nj = G.j1 - G.j0;
    // the group's stencils follow the knot box in LDS (5 float4 each)
    Stencil2D *sts = reinterpret_cast<Stencil2D *>(lds + LX * LY);
    __syncthreads();
    for (int k = threadIdx.x; k < LX * LY; k += blockDim.x) {
      const int lx = k / LY, ly = k - lx * LY;
      const int gx = x0 + G.bx0 + lx, gy = y0 + G.by0 + ly;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (gx >= 0 && gx < nx && gy >= 0 && gy < ny) v = ev.planes[(long)gx * ny + gy];
      lds[k] = v;
    }
    {
      const float4 *src = reinterpret_cast<const float4 *>(ev.st + G.j0);
      float4 *dst = reinterpret_cast<float4 *>(sts);
      for (int k = threadIdx.x; k < nj * 5; k += blockDim.x) dst[k] = src[k];
    }
    __syncthreads();
    // Samples of a group are sorted by knot offset on the host, so the knot
    // columns are read again only when the offset changes.
    int cur_ox = 1 << 30, cur_oy = 1 << 30, base = 0;
    float4 K0[M + 1], K1[M + 1];
    for (int jj = js; jj < nj; jj += JS) {
      const Stencil2D S = sts[jj];
      const int ox = __builtin_amdgcn_readfirstlane(S.ox), oy = __builtin_amdgcn_readfirstlane(S.oy);
      if (ox != cur_ox || oy != cur_oy) {
        cur_ox = ox;
        cur_oy = oy;
        base = (xr + ox - G.bx0) * LY + (yl + oy - G.by0);
#pragma unroll
        for (int a = 0; a <= M; ++a) {
          K0[a] = lds[base + a * LY];
          K1[a] = lds[base + a * LY + 1];
        }
      }
      float Zy[M + 1], ZXy[M + 1];
#pragma unroll
      for (int a = 0; a <= M; ++a) {
        const float4 k0 = K0[a], k1 = K1[a];
        Zy[a] = fmaf(S.hy[0], k0.x, fmaf(S.hy[1], k0.y, fmaf(S.hy[2], k1.x, S.hy[3] * k1.y)));
        ZXy[a] = fmaf(S.hy[0], k0.z, fmaf(S.hy[1], k0.w, fmaf(S.hy[2], k1.z, S.hy[3] * k1.w)));
      }
      float v[M];
#pragma unroll
      for (int m = 0; m < M; ++m)
        v[m] = fmaf(S.hx[0], Zy[m], fmaf(S.hx[1], ZXy[m], fmaf(S.hx[2], Zy[m + 1], S.hx[3] * ZXy[m + 1])));
      float vmin = v[0];
#pragma unroll
      for (int m = 1; m < M; ++m) vmin = fminf(vmin, v[m]);
      // src/FreeStream.cpp:245-246: bilinear where the spline is not > 0
      if (__ballot(!(vmin > 0.0f))) {
        float Bl[M + 1];
#pragma unroll
        for (int a = 0; a <= M; ++a) {
          const float z0 = K0[a].x, z1 = K1[a].x;
          Bl[a] = fmaf(S.ty, z1 - z0, z0);
        }
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const float bl = fmaf(S.tx, Bl[m + 1] - Bl[m], Bl[m]);
          v[m] = v[m] > 0.0f ? v[m] : bl;
        }
      }
      if (!interior) {
        const bool yok = mask_ok2(iy, iy + oy, S.dyj, ev.ymin, ev.dy, ny);
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const bool ok = yok && mask_ok2(ixb + m, ixb + m + ox, S.dxj, ev.xmin, ev.dx, nx);
          v[m] = ok ? v[m] : 0.0f;
        }
      }
#pragma unroll
      for (int m = 0; m < M; ++m) {
        acc[m][0] += v[m];
        acc[m][1] = fmaf(v[m], S.c, acc[m][1]);
        acc[m][2] = fmaf(v[m], S.s, acc[m][2]);
        acc[m][3] = fmaf(v[m], S.cc, acc[m][3]);
        acc[m][4] = fmaf(v[m], S.cs, acc[m][4]);
      }
    }
  }
  if (JS > 1) {
    // reduce the JS partial sums of each cell run through LDS
    float *red = reinterpret_cast<float *>(lds);
    __syncthreads();
    if (js != 0) {
#pragma unroll
      for (int m = 0; m < M; ++m)
#pragma unroll
        for (int c = 0; c < 5; ++c) red[((wave * 64 + lane) * M + m) * 5 + c] = acc[m][c];
    }
    __syncthreads();
    if (js == 0) {
#pragma unroll
      for (int w = 1; w < JS; ++w)
#pragma unroll
        for (int m = 0; m < M; ++m)
#pragma unroll
          for (int c = 0; c < 5; ++c) acc[m][c] += red[(((wave + w) * 64 + lane) * M + m) * 5 + c];
    }
  }
  if (js != 0) return;  // after the last barrier; the one-event kernels have no loop
  // src/LandauMatch.cpp:104: T = sum * d_phip / TAU.  The s^2 row is
  // sum(v) - sum(v c^2) (c^2 + s^2 = 1); rows 3, 6, 8, 9 carry tanh(0) = 0.
  const long P = ev.t_plane;
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int ix = ixb + m;
    if (ix < nx && iy < ny) {
      const long is = (long)ix * ny + iy;
      float r[6];
      const float s5 = acc[m][0] - acc[m][3];
#pragma unroll
      for (int c = 0; c < 5; ++c) r[c] = ((acc[m][c] * ev.out_scale) * ev.d_phip) / ev.tau;
      r[5] = ((s5 * ev.out_scale) * ev.d_phip) / ev.tau;
      ev.T[0 * P + is] = r[0];
      ev.T[1 * P + is] = r[1];
      ev.T[2 * P + is] = r[2];
      ev.T[3 * P + is] = 0.0f;
      ev.T[4 * P + is] = r[3];
      ev.T[5 * P + is] = r[4];
      ev.T[6 * P + is] = 0.0f;
      ev.T[7 * P + is] = r[5];
      ev.T[8 * P + is] = 0.0f;
      ev.T[9 * P + is] = 0.0f;
    }
  }
}

template <int M, int JS>
__global__ __launch_bounds__(256) void k_stream2d_one(Event2D ev) {
  extern __shared__ float4 lds4[];
  stream2d_tile_v2<M, JS>(ev, blockIdx.x % ev.tiles_x, blockIdx.x / ev.tiles_x, lds4);
}

// Batched: item = (event, tile x, tile y); each event keeps the tile shape
// (variant) its plan chose, so an event gives the same bits alone and in a
// batch.  The variant is uniform per item, so the switch does not diverge.
__global__ __launch_bounds__(256) void k_stream2d_batch(const Event2D *__restrict__ evs,
                                                        const int4 *__restrict__ items, int n_items) {
  extern __shared__ float4 lds4[];
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const int4 item = items[it];
    const Event2D ev = evs[item.x];
    switch (ev.variant) {
      case 1: stream2d_tile_v2<4, 4>(ev, item.y, item.z, lds4); break;
      case 2: stream2d_tile_v2<2, 4>(ev, item.y, item.z, lds4); break;
      default: stream2d_tile_v2<4, 1>(ev, item.y, item.z, lds4); break;
    }
    __syncthreads();  // all waves leave the tile before the next item reuses LDS
  }
}

}  // namespace

// Variant table: {M, JS}.  Variant 0 is the batch/throughput shape.
int fsm_stream2d_tile_w(int variant) {
  switch (variant) {
    case 1: return 4 / 4 * 2 * 4;  // M=4, JS=4
    case 2: return 4 / 4 * 2 * 2;  // M=2, JS=4
    default: return 4 * 2 * 4;     // M=4, JS=1
  }
}
int fsm_stream2d_tile_h(int) { return 32; }

size_t fsm_stream2d_lds_bytes(int variant, int box_floats4) {
  size_t box = (size_t)box_floats4 * sizeof(float4);
  size_t red = 0;
  if (variant == 1) red = (size_t)256 * 4 * 5 * sizeof(float);
  if (variant == 2) red = (size_t)256 * 2 * 5 * sizeof(float);
  return box > red ? box : red;
}

hipError_t fsm_launch_stream2d(const Event2D &ev, hipStream_t s) {
  const int tiles = ev.tiles_x * ev.tiles_y;
  const size_t lds = fsm_stream2d_lds_bytes(ev.variant, ev.lds_floats4);
  switch (ev.variant) {
    case 1: hipLaunchKernelGGL((k_stream2d_one<4, 4>), dim3(tiles), dim3(256), lds, s, ev); break;
    case 2: hipLaunchKernelGGL((k_stream2d_one<2, 4>), dim3(tiles), dim3(256), lds, s, ev); break;
    default: hipLaunchKernelGGL((k_stream2d_one<4, 1>), dim3(tiles), dim3(256), lds, s, ev); break;
  }
  return hipGetLastError();
}

hipError_t fsm_launch_stream2d_batch(const Event2D *evs, const int4 *items, int n, size_t lds_bytes, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int grid = n < 16384 ? n : 16384;
  hipLaunchKernelGGL(k_stream2d_batch, dim3(grid), dim3(256), lds_bytes, s, evs, items, n);
  return hipGetLastError();
}

hipError_t fsm_stream2d_set_lds_limit(int bytes) {
  const void *fns[] = {(const void *)k_stream2d_one<4, 1>, (const void *)k_stream2d_one<4, 4>,
                       (const void *)k_stream2d_one<2, 4>, (const void *)k_stream2d_batch};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
