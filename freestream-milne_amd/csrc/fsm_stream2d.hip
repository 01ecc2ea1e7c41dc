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

#include <algorithm>

#include "fsm_internal.h"
#include "fsm_kernels.h"

namespace {

// the 3x3 Landau frame and outputs of the fused epilogue (Event2D::fuse_landau)
#include "fsm_landau.h"

// The same as a call with scalar arguments, for the FP32 tiles that an MFMA
// kernel runs where the spline can undershoot (rare): inlined there too, the
// MFMA kernel's code would exceed its register cap of 168 and spill.
__device__ __noinline__ void landau2d_store_call(double *out, long out_plane, int eos, int viscous, int fill,
                                                 float tau, double inv_tau2, long is, float t0, float t1, float t2,
                                                 float t3, float t4, float t5) {
  Event2D s;
  s.l_out = out;
  s.l_out_plane = out_plane;
  s.l_eos = eos;
  s.l_viscous = viscous;
  s.l_fill = fill;
  s.tau = tau;
  s.l_inv_tau2 = inv_tau2;
  const float t6[6] = {t0, t1, t2, t3, t4, t5};
  landau2d_cell_store(s, t6, is);
}


// Copies the knot box [bx, bx+LX) x [by, by+LY) of an (nx, ny) x-major plane
// into LDS (zero outside the grid), column pitch `pitch` >= LY elements.  Each
// thread keeps U loads in flight and walks the box with an incremental (row,
// column) instead of a division.
template <int U, typename T>
__device__ __forceinline__ void stage_box(T *dst, const T *__restrict__ src, int bx, int by, int LX, int LY, int nx,
                                          int ny, int pitch) {
  const int n = LX * LY, step = blockDim.x;
  const int dlx = step / LY, dly = step - dlx * LY;
  int lx = threadIdx.x / LY, ly = threadIdx.x - lx * LY;
  for (int k0 = threadIdx.x; k0 < n; k0 += U * step) {
    T v[U];
    int di[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int kx = bx + lx, ky = by + ly;
      di[u] = k0 + u * step < n ? lx * pitch + ly : -1;
      v[u] = T{};
      if (di[u] >= 0 && (unsigned)kx < (unsigned)nx && (unsigned)ky < (unsigned)ny) v[u] = src[(long)kx * ny + ky];
      lx += dlx;
      ly += dly;
      if (ly >= LY) {
        ly -= LY;
        ++lx;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (di[u] >= 0) dst[di[u]] = v[u];
  }
}

// Copies n 16-byte elements with U loads in flight per thread.
template <int U, typename T>
__device__ __forceinline__ void stage_copy(T *dst, const T *__restrict__ src, int n) {
  for (int k0 = threadIdx.x; k0 < n; k0 += U * blockDim.x) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u * blockDim.x;
      if (k < n) v[u] = src[k];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u * blockDim.x;
      if (k < n) dst[k] = v[u];
    }
  }
}

// The reference's value of a kept sample whose source point is not strictly
// inside the spline's knot range, where it uses bilinear interpolation with
// its own float fractions (src/FreeStream.cpp:198-246; the range test compares
// the float source point, promoted to double, with the grid coordinates).
// For a kept sample that only happens on the last knot line: a shift smaller
// than half an ulp of the coordinate leaves x_new == x_max, and the rounded
// (x_new - xmin)/DX can fall just below n - 1, so the reference keeps the
// sample at floor n - 2 with a fraction near 1.  Returns false (nothing
// written) when the point is inside the range.  Call only for kept samples.
__device__ __forceinline__ bool ref_off_spline(const Event2D &ev, const Stencil2D &S, int ix, int iy, float *out) {
#pragma clang fp contract(off)
  const float x = (float)ix * ev.dx + ev.xmin, y = (float)iy * ev.dy + ev.ymin;
  const float xn = x - S.dxj, yn = y - S.dyj;
  const double xmax = (double)((float)(ev.nx - 1) * ev.dx + ev.xmin);
  const double ymax = (double)((float)(ev.ny - 1) * ev.dy + ev.ymin);
  if ((double)ev.xmin < (double)xn && (double)xn < xmax && (double)ev.ymin < (double)yn && (double)yn < ymax)
    return false;
  const float ixn = (xn - ev.xmin) / ev.dx, iyn = (yn - ev.ymin) / ev.dy;
  const float fx = floorf(ixn), fy = floorf(iyn);
  if (!(fx >= 1.0f && fx < (float)(ev.nx - 1) && fy >= 1.0f && fy < (float)(ev.ny - 1))) return false;
  const float x0 = ixn - fx, x1 = iyn - fy;
  const int i0 = (int)fx, j0 = (int)fy;
  const float a00 = fsm_gload(ev.planes + (long)i0 * ev.ny + j0).x;
  const float a10 = fsm_gload(ev.planes + (long)(i0 + 1) * ev.ny + j0).x;
  const float a01 = fsm_gload(ev.planes + (long)i0 * ev.ny + j0 + 1).x;
  const float a11 = fsm_gload(ev.planes + (long)(i0 + 1) * ev.ny + j0 + 1).x;
  *out = ((1 - x0) * (1 - x1) * a00) + ((x0) * (1 - x1) * a10) + ((1 - x0) * (x1) * a01) + ((x0) * (x1) * a11);
  return true;
}

template <int M, int JS, bool OUTLINE_LANDAU = false>
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
    stage_box<4>(lds, ev.planes, x0 + G.bx0, y0 + G.by0, LX, LY, nx, ny, LY);
    stage_copy<4>(reinterpret_cast<float4 *>(sts), reinterpret_cast<const float4 *>(ev.st + G.j0), nj * 5);
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
        const bool yok = fsm_mask_ok(iy, iy + oy, S.dyj, ev.ymin, ev.dy, ny);
        const bool ymax_line = iy + oy >= ny - 2;
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const bool ok = yok && fsm_mask_ok(ixb + m, ixb + m + ox, S.dxj, ev.xmin, ev.dx, nx);
          v[m] = ok ? v[m] : 0.0f;
          // a kept sample on the last knot line: the reference's bilinear
          if (ok && (ymax_line || ixb + m + ox >= nx - 2)) (void)ref_off_spline(ev, S, ixb + m, iy, &v[m]);
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
  if (ev.fuse_landau) {
    // Fused Landau stage: the lane's M cells one at a time, their moments
    // staged in the lane's own LDS slot (M * 5 floats: the slot the JS
    // reduction gives this wave, which no other wave reads), so that the
    // solve of one cell is not interleaved with the others' (registers).
    float *slot = reinterpret_cast<float *>(lds) + (wave * 64 + lane) * (M * 5);
    if (JS == 1) __syncthreads();  // every wave is done with the last group's knot box
#pragma unroll
    for (int m = 0; m < M; ++m)
#pragma unroll
      for (int c = 0; c < 5; ++c) slot[m * 5 + c] = acc[m][c];
#pragma unroll 1
    for (int m = 0; m < M; ++m) {
      const int ix = ixb + m;
      if (ix >= nx || iy >= ny) continue;
      float a[5], r[6];
#pragma unroll
      for (int c = 0; c < 5; ++c) a[c] = slot[m * 5 + c];
      const float s5 = a[0] - a[3];
#pragma unroll
      for (int c = 0; c < 5; ++c) r[c] = ((a[c] * ev.out_scale) * ev.d_phip) / ev.tau;
      r[5] = ((s5 * ev.out_scale) * ev.d_phip) / ev.tau;
      // the Landau stage on these floats
      if (OUTLINE_LANDAU)
        landau2d_store_call(ev.l_out, ev.l_out_plane, ev.l_eos, ev.l_viscous, ev.l_fill, ev.tau, ev.l_inv_tau2,
                            (long)ix * ny + iy, r[0], r[1], r[2], r[3], r[4], r[5]);
      else
        landau2d_cell_store(ev, r, (long)ix * ny + iy);
    }
    return;
  }
  // src/LandauMatch.cpp:104: T = sum * d_phip / TAU.  The s^2 row is
  // sum(v) - sum(v c^2) (c^2 + s^2 = 1); rows 3, 6, 8, 9 carry tanh(0) = 0.
  const long P = ev.t_plane;
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int ix = ixb + m;
    if (ix < nx && iy < ny && ev.out_j) {
      // src/LandauMatch.cpp:140: J = sum * DRAP * d_phip; rows 1, c, s, tanh(0) = 0
      const long is = (long)ix * ev.t_row + iy;
#pragma unroll
      for (int c = 0; c < 3; ++c) ev.T[c * P + is] = ((acc[m][c] * ev.out_scale) * ev.jscale) * ev.d_phip;
      ev.T[3 * P + is] = 0.0f;
    } else if (ix < nx && iy < ny) {
      const long is = (long)ix * ev.t_row + iy;
      float r[6];
      const float s5 = acc[m][0] - acc[m][3];
#pragma unroll
      for (int c = 0; c < 5; ++c) r[c] = ((acc[m][c] * ev.out_scale) * ev.d_phip) / ev.tau;
      r[5] = ((s5 * ev.out_scale) * ev.d_phip) / ev.tau;
      ev.T[0 * P + is] = r[0];
      ev.T[1 * P + is] = r[1];
      ev.T[2 * P + is] = r[2];
      ev.T[4 * P + is] = r[3];
      ev.T[5 * P + is] = r[4];
      ev.T[7 * P + is] = r[5];
      if (!ev.skip_eta_planes) {
        ev.T[3 * P + is] = 0.0f;
        ev.T[6 * P + is] = 0.0f;
        ev.T[8 * P + is] = 0.0f;
        ev.T[9 * P + is] = 0.0f;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Mixed-precision tile: the knot-offset form of the quadrature on FP16 MFMA
// (the algebra is described above FSM_MFMA_ZTARGET in fsm_internal.h).
//
// v_mfma_f32_16x16x32_f16 operand layout (wave64): lane l holds A[l%16][8*(l/16)+0..7],
// B[8*(l/16)+0..7][l%16], and D[4*(l/16)+0..3][l%16].  Rows of A are 16 cells along y,
// the K index is 8*corner + (hi z, zy, zx, zxy, lo z, zy, zx, zxy), so lane
// (corner, cell) reads its knot's 16-byte record from LDS as one load; columns
// 0-5 of B carry W_h and columns 8-13 W_l, each for both knot halves, so
// D[:, m] + D[:, m+8] = (K_h + K_l)(W_h + W_l) for moment m
// (1, c, s, c^2, cs, s^2).
//
// Wave w owns tile columns [w*XW, (w+1)*XW) x 32 rows (two 16-cell blocks).
// Samples whose shifted floor index is not the same for every cell (the
// stencil offset is rounding-dependent, e.g. sin(0) = 0) are in W too; for
// cells where that makes the edge mask rounding-dependent a VALU pass applies
// the reference's exact decision (sample_correction).
typedef _Float16 fsm_half8 __attribute__((ext_vector_type(8)));
typedef float fsm_float4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 plane_or_zero(const Event2D &ev, int gx, int gy) {
  if ((unsigned)gx < (unsigned)ev.nx && (unsigned)gy < (unsigned)ev.ny) return fsm_gload(ev.planes + (long)gx * ev.ny + gy);
  return make_float4(0.f, 0.f, 0.f, 0.f);
}

// Edge correction of one rounding-dependent sample of one cell.  The MFMA
// form adds the bicubic value at the stencil offset whenever the offset's
// floor index ix + ox is in [1, n-2].  For a sample whose shift is within
// rounding of an integer the reference's floor can be ix + ox -+ 1; away from
// the mask boundary that does not change the decision (and the value is
// continuous across the knot line), but for a floor index in {0, 1, n-2, n-1}
// the reference's own float expression decides (fsm_mask_ok).  Returns the exact
// sample (bilinear where the bicubic is not > 0, src/FreeStream.cpp:245-246)
// minus what the MFMA form added.
__device__ __forceinline__ float sample_correction(const Event2D &ev, const Stencil2D &S, int ix, int iy) {
  const int gx = ix + S.ox, gy = iy + S.oy;
  const bool bx = (unsigned)gx <= 1u || (unsigned)(gx - (ev.nx - 2)) <= 1u;
  const bool by = (unsigned)gy <= 1u || (unsigned)(gy - (ev.ny - 2)) <= 1u;
  if (!bx && !by) return 0.0f;
  const bool idx = (unsigned)(gx - 1) <= (unsigned)(ev.nx - 3) && (unsigned)(gy - 1) <= (unsigned)(ev.ny - 3);
  const bool ok = fsm_mask_ok(iy, gy, S.dyj, ev.ymin, ev.dy, ev.ny) && fsm_mask_ok(ix, gx, S.dxj, ev.xmin, ev.dx, ev.nx);
  if (!ok && !idx) return 0.0f;
  const float4 k00 = plane_or_zero(ev, gx, gy), k01 = plane_or_zero(ev, gx, gy + 1);
  const float4 k10 = plane_or_zero(ev, gx + 1, gy), k11 = plane_or_zero(ev, gx + 1, gy + 1);
  const float zy0 = fmaf(S.hy[0], k00.x, fmaf(S.hy[1], k00.y, fmaf(S.hy[2], k01.x, S.hy[3] * k01.y)));
  const float zxy0 = fmaf(S.hy[0], k00.z, fmaf(S.hy[1], k00.w, fmaf(S.hy[2], k01.z, S.hy[3] * k01.w)));
  const float zy1 = fmaf(S.hy[0], k10.x, fmaf(S.hy[1], k10.y, fmaf(S.hy[2], k11.x, S.hy[3] * k11.y)));
  const float zxy1 = fmaf(S.hy[0], k10.z, fmaf(S.hy[1], k10.w, fmaf(S.hy[2], k11.z, S.hy[3] * k11.w)));
  const float v = fmaf(S.hx[0], zy0, fmaf(S.hx[1], zxy0, fmaf(S.hx[2], zy1, S.hx[3] * zxy1)));
  float exact = v;
  if (!(v > 0.0f)) {
    const float b0 = fmaf(S.ty, k01.x - k00.x, k00.x), b1 = fmaf(S.ty, k11.x - k10.x, k10.x);
    exact = fmaf(S.tx, b1 - b0, b0);
  }
  if (ok) (void)ref_off_spline(ev, S, ix, iy, &exact);
  return (ok ? exact : 0.0f) - (idx ? v : 0.0f);
}

// B operands are stored packed: lane l's four distinct halves (8 bytes, W_h
// in columns 0-5, W_l in columns 8-13, zero in the others), repeated in
// halves 4-7 so that both knot halves meet both weight halves.  `b` is the
// operand for cell columns sharing the lane's corner roles, `bs` for the odd
// columns (B'[lane] = B[lane ^ 16]: ds_swizzle bit mode and 0x1f, xor 0x10).
__device__ __forceinline__ void bmat_expand(uint2 v, fsm_half8 &b, fsm_half8 &bs) {
  const uint2 w = make_uint2((unsigned)__builtin_amdgcn_ds_swizzle((int)v.x, 0x401F),
                             (unsigned)__builtin_amdgcn_ds_swizzle((int)v.y, 0x401F));
  b = __builtin_bit_cast(fsm_half8, make_uint4(v.x, v.y, v.x, v.y));
  bs = __builtin_bit_cast(fsm_half8, make_uint4(w.x, w.y, w.x, w.y));
}

// The packed B operands of a group's MFMAs stay in global memory (L2: every
// tile of an event reads the same few tens of KiB); lane l's 8 bytes of
// MFMA m are bg[64 m + l].  They are loaded two MFMA records ahead of their
// use, so the loads are in flight during the previous records' MFMAs, and
// they take no LDS, which leaves the group's LDS image to its knot box.
struct BPrefetch {
  const uint2 *bg;
  uint2 v0, v1;
  int n;
  __device__ __forceinline__ BPrefetch(const uint2 *bg_, int n_, int lane) : bg(bg_ + lane), n(n_) {
    v0 = n > 0 ? fsm_gload(bg) : make_uint2(0u, 0u);
    v1 = n > 1 ? fsm_gload(bg + 64) : make_uint2(0u, 0u);
  }
  // B of record o (records are taken in order); starts the load of o + 2
  __device__ __forceinline__ uint2 next(int o) {
    const uint2 v = v0;
    v0 = v1;
    if (o + 2 < n) v1 = fsm_gload(bg + 64 * (o + 2));
    return v;
  }
};

// The MFMA loop over the knot offsets of one LDS stage, for the wave's XW
// cell columns x 2 row blocks.  Adjacent cell columns share a knot column:
// register R[k] holds knot column 2k (lanes with corner-x 0) and 2k+1 (corner
// x 1), which is the operand of cell column 2k with weights B; cell column
// 2k+1 takes knot column 2k+1 from R[k] and 2k+2 from R[k+1] (a lane select)
// with weights B' (corner-x roles swapped: B'[lane] = B[lane ^ 16], a lane
// swizzle).  So a wave loads XW/2 + 1 knot-column pairs per row block instead
// of XW.
// MASKED: edge mask (src/FreeStream.cpp:213), exact here since the floor
// index of an MFMA sample is ix + ox for every cell.  Rows out of range zero
// their knot registers (a 16-row block is all in, all out, or split), columns
// out of range skip their MFMAs; both tests are wave-uniform.  A second row
// block with no live cell (y0 + 16 >= ny) is skipped in both forms.
template <int XW, bool MASKED>
__device__ __forceinline__ void mfma_offsets(fsm_float4v (&acc)[XW][2], const uint4 *lh, const uint2 *bg,
                                             const int4 *orec, int no, int lane_base, int LYP, int gx, int r,
                                             int xw0, int y0, int nx, int ny) {
  const int lane = threadIdx.x & 63;
  constexpr int NR = XW / 2 + 1;
  const int nyb = y0 + 16 < ny ? 2 : 1;  // row blocks holding live cells
  no = __builtin_amdgcn_readfirstlane(no);
  BPrefetch bq(bg, no, lane);
  for (int o = 0; o < no; ++o) {
    const int4 rec = orec[o];
    const int ox = __builtin_amdgcn_readfirstlane(rec.x), oy = __builtin_amdgcn_readfirstlane(rec.y);
    const int delta = __builtin_amdgcn_readfirstlane(rec.z);
    fsm_half8 b0, b1;
    bmat_expand(bq.next(o), b0, b1);
#pragma unroll
    for (int yb = 0; yb < 2; ++yb) {
      if (yb >= nyb) break;
      // one 16-row block at a time: the knot registers of a block are
      // reused by its XW columns and then released
      uint4 R[NR][1];
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        // the last pair is used only by corner-x 1 lanes of cell column XW-1
        // through R[XW/2] at corner-x 0, so it stays inside the box
        const int col = (k < XW / 2 || XW == 1) ? lane_base : lane_base - gx * LYP;
        R[k][0] = lh[col + delta + 2 * k * LYP + yb * 16];
      }
      if (MASKED) {
        const int ylo = y0 + yb * 16 + oy;  // floor index of row r is ylo + r
        if (ylo + 15 < 1 || ylo > ny - 2) continue;
        if (!(ylo >= 1 && ylo + 15 <= ny - 2)) {
          const bool yok = (unsigned)(ylo + r - 1) <= (unsigned)(ny - 3);
#pragma unroll
          for (int k = 0; k < NR; ++k) {
            R[k][0].x = yok ? R[k][0].x : 0u;
            R[k][0].y = yok ? R[k][0].y : 0u;
            R[k][0].z = yok ? R[k][0].z : 0u;
            R[k][0].w = yok ? R[k][0].w : 0u;
          }
        }
      }
#pragma unroll
      for (int xc = 0; xc < XW; ++xc) {
        if (MASKED && (unsigned)(xw0 + xc + ox - 1) > (unsigned)(nx - 3)) continue;
        uint4 a = R[xc / 2][0];
        if (xc & 1) {
          const uint4 n1 = R[xc / 2 + 1][0];
          a.x = gx ? a.x : n1.x;
          a.y = gx ? a.y : n1.y;
          a.z = gx ? a.z : n1.z;
          a.w = gx ? a.w : n1.w;
        }
        acc[xc][yb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(fsm_half8, a), (xc & 1) ? b1 : b0,
                                                             acc[xc][yb], 0, 0, 0);
      }
    }
  }
}

// Interior form of one LDS stage: the group's knot displacements (offset +
// corner, weights summed per displacement) covered by dominoes (dx, dx+1) at
// one row dy; one MFMA takes two dominoes, lane group (gx, gy) reading knot
// (x + dx_gy + gx, y + dy_gy), so only the row offset differs between the gy
// halves (a per-lane box delta).  The column sharing of mfma_offsets is
// unchanged.  Used where every cell of the wave keeps every sample of the
// group (no masks), so merging samples of different offsets is exact.
template <int XW>
__device__ __forceinline__ void mfma_pairs(fsm_float4v (&acc)[XW][2], const uint4 *lh, const uint2 *bg,
                                           const int4 *prec, int np, int lane_base, int LYP, int gx, int gy,
                                           int nyb) {
  const int lane = threadIdx.x & 63;
  constexpr int NR = XW / 2 + 1;
  np = __builtin_amdgcn_readfirstlane(np);
  BPrefetch bq(bg, np, lane);
  for (int o = 0; o < np; ++o) {
    const int4 rec = prec[o];
    const int delta = gy ? rec.y : rec.x;
    fsm_half8 b0, b1;
    bmat_expand(bq.next(o), b0, b1);
#pragma unroll
    for (int yb = 0; yb < 2; ++yb) {
      if (yb >= nyb) break;
      uint4 R[NR];
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        const int col = (k < XW / 2 || XW == 1) ? lane_base : lane_base - gx * LYP;
        R[k] = lh[col + delta + 2 * k * LYP + yb * 16];
      }
#pragma unroll
      for (int xc = 0; xc < XW; ++xc) {
        uint4 a = R[xc / 2];
        if (xc & 1) {
          const uint4 n1 = R[xc / 2 + 1];
          a.x = gx ? a.x : n1.x;
          a.y = gx ? a.y : n1.y;
          a.z = gx ? a.z : n1.z;
          a.w = gx ? a.w : n1.w;
        }
        acc[xc][yb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(fsm_half8, a), (xc & 1) ? b1 : b0,
                                                             acc[xc][yb], 0, 0, 0);
      }
    }
  }
}

// A zero knot in global memory: the source of box slots off the grid and of
// the padding rows of each box column.
__device__ const uint4 fsm_zero_knot = {0u, 0u, 0u, 0u};

// LDS image of one offset group, copied straight from global memory (16 B
// per lane, a wave fills 64 consecutive slots): the knot box [LX][LYP]
// (knots off the grid and rows LY..LYP-1 read as zero), then the records of
// the group's MFMAs of both forms.  Wave w fills chunks w, w + 4, ... of each
// part.
__device__ __forceinline__ void group_dma(const Event2D &ev, const StencilGroup &G, uint4 *lh, int x0, int y0,
                                          int TX, int TY, int wave, int lane) {
  const int LX = TX + G.bx1 - G.bx0, LY = TY + G.by1 - G.by0, LYP = (LY + 15) & ~15;
  const int n = LX * LYP;
  const int no = G.j1 - G.j0 + G.pad1;
  const int bx = x0 + G.bx0, by = y0 + G.by0;
  // slot s = 64 c + lane at box (lx, ly); s advances by 256 per chunk step
  const int s0 = wave * 64 + lane;
  int lx = s0 / LYP, ly = s0 - lx * LYP;
  const int dlx = 256 / LYP, dly = 256 - dlx * LYP;
  for (int c = wave; c * 64 < n; c += 4) {
    const int kx = bx + lx, ky = by + ly;
    const bool in = ly < LY && (unsigned)kx < (unsigned)ev.nx && (unsigned)ky < (unsigned)ev.ny;
    const uint4 *src = in ? ev.planes_h + (kx * ev.ny + ky) : &fsm_zero_knot;
    if (c * 64 + lane < n) fsm_lds_dma16(src, lh + c * 64);
    lx += dlx;
    ly += dly;
    if (ly >= LYP) {
      ly -= LYP;
      ++lx;
    }
  }
  uint4 *rl = lh + n;
  const uint4 *rsrc = reinterpret_cast<const uint4 *>(ev.moffs + G.pad0);
  for (int c = wave; c * 64 < no; c += 4)
    if (c * 64 + lane < no) fsm_lds_dma16(rsrc + c * 64 + lane, rl + c * 64);
}

// MFMA tile shapes and the FP32 tile of the same width they fall back to.
template <int V> struct MfmaShape;
template <> struct MfmaShape<0> { static constexpr int XW = 8, M = 4, JS = 1; };
template <> struct MfmaShape<1> { static constexpr int XW = 2, M = 4, JS = 4; };
template <> struct MfmaShape<2> { static constexpr int XW = 1, M = 2, JS = 4; };

template <int V>
__device__ void stream2d_mfma_tile(const Event2D &ev, int tile_x, int tile_y, float4 *lds) {
  constexpr int XW = MfmaShape<V>::XW;
  constexpr int TX = 4 * XW, TY = 32, NC = TX * TY;
  const int nx = ev.nx, ny = ev.ny;
  const int x0 = tile_x * TX, y0 = tile_y * TY;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  {
    // spline tiles holding the knots this tile can read
    const int sx0 = max(0, x0 + ev.rbx0) / FSMSP_TX, sx1 = min(nx - 1, x0 + TX - 1 + ev.rbx1) / FSMSP_TX;
    const int sy0 = max(0, y0 + ev.rby0) / FSMSP_TY, sy1 = min(ny - 1, y0 + TY - 1 + ev.rby1) / FSMSP_TY;
    const int cy = sy1 - sy0 + 1, cnt = (sx1 - sx0 + 1) * cy;
    int bad = 0;
    for (int k = tid; k < cnt; k += blockDim.x)
      bad |= fsm_gload(ev.sp_flags + (sx0 + k / cy) * ev.sp_tiles_y + sy0 + k % cy) != 0u;
    if (__syncthreads_or(bad)) {
      stream2d_tile_v2<MfmaShape<V>::M, MfmaShape<V>::JS, true>(ev, tile_x, tile_y, lds);
      return;
    }
  }
  const int g = lane >> 4, r = lane & 15, gx = g & 1, gy = g >> 1;
  uint4 *lh = reinterpret_cast<uint4 *>(lds);
  fsm_float4v acc[XW][2];
#pragma unroll
  for (int xc = 0; xc < XW; ++xc)
#pragma unroll
    for (int yb = 0; yb < 2; ++yb) acc[xc][yb] = fsm_float4v{0.f, 0.f, 0.f, 0.f};
  const int ylast = min(y0 + TY - 1, ny - 1);
  for (int gi = 0; gi < ev.n_mgroups; ++gi) {
    const StencilGroup G = ev.mgroups[gi];
    // box columns are padded to 16 records (256 B) so that the lanes reading
    // knot columns c and c+1 do not collide in the LDS banks
    const int LX = TX + G.bx1 - G.bx0, LY = TY + G.by1 - G.by0, LYP = (LY + 15) & ~15;
    const int no = G.j1 - G.j0, np = G.pad1;
    // packed B operands in global memory: offsets, then pair MFMAs
    const uint2 *bg = reinterpret_cast<const uint2 *>(ev.bmat + (long)G.pad0 * FSM_BMAT_U4);
    int4 *orec = reinterpret_cast<int4 *>(lh + LX * LYP);  // offsets (ox, oy, box delta, -), pairs (delta gy=0, gy=1, -, -)
    __syncthreads();  // every wave is done with the previous group's image
    group_dma(ev, G, lh, x0, y0, TX, TY, wave, lane);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this wave's copies have landed in LDS
    __syncthreads();
    // every real cell's floor index stays inside [1, n-2] for all offsets of the group
    // lane's knot (corner) for cell column wave*XW, row r, at box offset 0
    const int lane_base = (wave * XW + gx) * LYP + r + gy;
    // this wave's cells keep every floor index of the group inside [1, n-2]
    const int xw0 = x0 + wave * XW, xw1 = min(xw0 + XW - 1, nx - 1);
    const bool interior = (xw0 + G.bx0 >= 1) && (xw1 + G.bx1 - 1 <= nx - 2) && (y0 + G.by0 >= 1) &&
                          (ylast + G.by1 - 1 <= ny - 2);
    // waves whose columns hold no live cell (the last tile column) skip the group
    if (xw0 < nx) {
      if (interior)
        mfma_pairs<XW>(acc, lh, bg + no * 64, orec + no, np, lane_base - gy, LYP, gx, gy,
                       y0 + 16 < ny ? 2 : 1);
      else
        mfma_offsets<XW, true>(acc, lh, bg, orec, no, lane_base, LYP, gx, r, xw0, y0, nx, ny);
    }
  }
  // moments to LDS: Tt[m][lx*TY + ly]
  __syncthreads();
  float *Tt = reinterpret_cast<float *>(lds);
  {
    const int n = lane & 15, q = lane >> 4;
#pragma unroll
    for (int xc = 0; xc < XW; ++xc)
#pragma unroll
      for (int yb = 0; yb < 2; ++yb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float v = acc[xc][yb][i];
          const float w = __shfl_down(v, 8, 16);
          if (n < 6) Tt[n * NC + (wave * XW + xc) * TY + yb * 16 + q * 4 + i] = v + w;
        }
  }
  __syncthreads();
  // src/LandauMatch.cpp:104: T = sum * d_phip / TAU; rows 3, 6, 8, 9 carry tanh(0) = 0
  const long P = ev.t_plane;
  const float inv_sc = 1.0f / fsm_sload(ev.mscale + FSM_KS_PARTS);  // a power of two: exact
  constexpr int CPT = (NC + 255) / 256;     // cells per thread
  // cell q of this thread: c = tid + 256 q, column c / TY, row c % TY (TY =
  // 32: shifts, recomputed where used rather than held in registers)
  auto cell_x = [&](int q) { return x0 + (tid + 256 * q) / TY; };
  auto cell_y = [&](int q) { return y0 + (tid + 256 * q) % TY; };
  auto cell_live = [&](int q) { return tid + 256 * q < NC && cell_x(q) < nx && cell_y(q) < ny; };
  float t[CPT][6];
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    const int c = tid + 256 * q;
    const bool lv = cell_live(q);
#pragma unroll
    for (int m = 0; m < 6; ++m) t[q][m] = lv ? Tt[m * NC + c] * inv_sc : 0.0f;
  }
  // rounding-dependent samples next to the mask boundary: FP32, the
  // reference's masks (the cells of a thread are evaluated together)
  for (int a = 0; a < ev.n_amb; ++a) {
    const Stencil2D S = ev.st[ev.amb[a]];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      const float v = cell_live(q) ? sample_correction(ev, S, cell_x(q), cell_y(q)) : 0.0f;
      t[q][0] += v;
      t[q][1] = fmaf(v, S.c, t[q][1]);
      t[q][2] = fmaf(v, S.s, t[q][2]);
      t[q][3] = fmaf(v, S.cc, t[q][3]);
      t[q][4] = fmaf(v, S.cs, t[q][4]);
      t[q][5] = fmaf(v, S.ss, t[q][5]);
    }
  }
  // src/LandauMatch.cpp:104: T = sum * d_phip / TAU; rows 3, 6, 8, 9 carry tanh(0) = 0
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    if (!cell_live(q)) continue;
    if (ev.out_j) {
      // src/LandauMatch.cpp:140: J = sum * DRAP * d_phip; rows 1, c, s, tanh(0) = 0
      const long is = (long)cell_x(q) * ev.t_row + cell_y(q);
#pragma unroll
      for (int c = 0; c < 3; ++c) fsm_gstore(ev.T + c * P + is, ((t[q][c] * ev.out_scale) * ev.jscale) * ev.d_phip);
      fsm_gstore(ev.T + 3 * P + is, 0.0f);
      continue;
    }
    float rr[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) rr[m] = ((t[q][m] * ev.out_scale) * ev.d_phip) / ev.tau;
    if (ev.fuse_landau) {
      landau2d_cell_store(ev, rr, (long)cell_x(q) * ny + cell_y(q));  // the Landau stage on these floats
      continue;
    }
    const long is = (long)cell_x(q) * ev.t_row + cell_y(q);
    fsm_gstore(ev.T + 0 * P + is, rr[0]);
    fsm_gstore(ev.T + 1 * P + is, rr[1]);
    fsm_gstore(ev.T + 2 * P + is, rr[2]);
    fsm_gstore(ev.T + 4 * P + is, rr[3]);
    fsm_gstore(ev.T + 5 * P + is, rr[4]);
    fsm_gstore(ev.T + 7 * P + is, rr[5]);
    if (!ev.skip_eta_planes) {
      fsm_gstore(ev.T + 3 * P + is, 0.0f);
      fsm_gstore(ev.T + 6 * P + is, 0.0f);
      fsm_gstore(ev.T + 8 * P + is, 0.0f);
      fsm_gstore(ev.T + 9 * P + is, 0.0f);
    }
  }
}

template <int V>
__global__ __launch_bounds__(256) void k_stream2d_mfma_one(Event2D ev) {
  extern __shared__ float4 lds4[];
  const int t = xcd_order(blockIdx.x, gridDim.x);
  stream2d_mfma_tile<V>(ev, t % ev.tiles_x, t / ev.tiles_x, lds4);
}

template <int M, int JS>
__global__ __launch_bounds__(256) void k_stream2d_one(Event2D ev) {
  extern __shared__ float4 lds4[];
  const int t = xcd_order(blockIdx.x, gridDim.x);
  stream2d_tile_v2<M, JS>(ev, t % ev.tiles_x, t / ev.tiles_x, lds4);
}

// Batched: item = (event, tile x, tile y); each event keeps the tile shape
// (variant) its plan chose, so an event gives the same bits alone and in a
// batch.  The variant is uniform per item, so the switch does not diverge.
__global__ __launch_bounds__(256) void k_stream2d_batch(const Event2D *__restrict__ evs,
                                                        const int4 *__restrict__ items, int n_items) {
  extern __shared__ float4 lds4[];
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const int4 item = fsm_sload(items + xcd_order(it, n_items));
    const Event2D ev = fsm_sload(evs + item.x);
    switch (ev.variant) {
      case 1: stream2d_tile_v2<4, 4>(ev, item.y, item.z, lds4); break;
      case 2: stream2d_tile_v2<2, 4>(ev, item.y, item.z, lds4); break;
      default: stream2d_tile_v2<4, 1>(ev, item.y, item.z, lds4); break;
    }
    __syncthreads();  // all waves leave the tile before the next item reuses LDS
  }
}

// Batched MFMA form (every 2-D event of a batch runs the context's path).
// Three workgroups per CU (the LDS image of a group is up to 48 KiB; the
// kernel needs 142 registers per lane, under the cap of 168 that three
// workgroups leave); two per CU measured 28% slower.
__global__ __launch_bounds__(256, 3) void k_stream2d_mfma_batch(const Event2D *__restrict__ evs,
                                                             const int4 *__restrict__ items, int n_items) {
  extern __shared__ float4 lds4[];
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const int4 item = fsm_sload(items + xcd_order(it, n_items));
    const Event2D ev = fsm_sload(evs + item.x);
    switch (ev.variant) {
      case 1: stream2d_mfma_tile<1>(ev, item.y, item.z, lds4); break;
      case 2: stream2d_mfma_tile<2>(ev, item.y, item.z, lds4); break;
      default: stream2d_mfma_tile<0>(ev, item.y, item.z, lds4); break;
    }
    __syncthreads();
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
  // the JS reduction, and the fused Landau epilogue's per-lane slots: M * 5 floats per thread
  size_t red = (size_t)256 * 4 * 5 * sizeof(float);
  if (variant == 2) red = (size_t)256 * 2 * 5 * sizeof(float);
  return box > red ? box : red;
}

size_t fsm_stream2d_event_lds_bytes(const Event2D &ev) {
  size_t b = fsm_stream2d_lds_bytes(ev.variant, ev.lds_floats4);
  if (ev.mfma) {
    const size_t tile = (size_t)6 * fsm_stream2d_tile_w(ev.variant) * fsm_stream2d_tile_h(ev.variant) * sizeof(float);
    b = std::max(b, std::max(tile, (size_t)ev.mfma_lds4 * 16));
  }
  return b;
}

hipError_t fsm_launch_stream2d(const Event2D &ev, hipStream_t s) {
  const int tiles = ev.tiles_x * ev.tiles_y;
  const size_t lds = fsm_stream2d_event_lds_bytes(ev);
  if (ev.mfma) {
    switch (ev.variant) {
      case 1: hipLaunchKernelGGL((k_stream2d_mfma_one<1>), dim3(tiles), dim3(256), lds, s, ev); break;
      case 2: hipLaunchKernelGGL((k_stream2d_mfma_one<2>), dim3(tiles), dim3(256), lds, s, ev); break;
      default: hipLaunchKernelGGL((k_stream2d_mfma_one<0>), dim3(tiles), dim3(256), lds, s, ev); break;
    }
    return hipGetLastError();
  }
  switch (ev.variant) {
    case 1: hipLaunchKernelGGL((k_stream2d_one<4, 4>), dim3(tiles), dim3(256), lds, s, ev); break;
    case 2: hipLaunchKernelGGL((k_stream2d_one<2, 4>), dim3(tiles), dim3(256), lds, s, ev); break;
    default: hipLaunchKernelGGL((k_stream2d_one<4, 1>), dim3(tiles), dim3(256), lds, s, ev); break;
  }
  return hipGetLastError();
}

hipError_t fsm_launch_stream2d_batch(const Event2D *evs, const int4 *items, int n, size_t lds_bytes, int mfma,
                                     hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int grid = n < 16384 ? n : 16384;
  if (mfma)
    hipLaunchKernelGGL(k_stream2d_mfma_batch, dim3(grid), dim3(256), lds_bytes, s, evs, items, n);
  else
    hipLaunchKernelGGL(k_stream2d_batch, dim3(grid), dim3(256), lds_bytes, s, evs, items, n);
  return hipGetLastError();
}

hipError_t fsm_stream2d_set_lds_limit(int bytes) {
  const void *fns[] = {(const void *)k_stream2d_one<4, 1>,   (const void *)k_stream2d_one<4, 4>,
                       (const void *)k_stream2d_one<2, 4>,   (const void *)k_stream2d_batch,
                       (const void *)k_stream2d_mfma_one<0>, (const void *)k_stream2d_mfma_one<1>,
                       (const void *)k_stream2d_mfma_one<2>, (const void *)k_stream2d_mfma_batch};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
