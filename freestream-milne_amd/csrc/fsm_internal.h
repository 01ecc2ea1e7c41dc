// fsm_internal.h -- structures shared by the host planner and the HIP kernels.
//
// Device data layout (HBM, all structure-of-arrays):
//   * 2-D spline planes: float4 per knot {z, z_y*hy, z_x*hx, z_xy*hx*hy},
//     x-major (knot index ix*NY + iy), z = converted density e0/(2 pi).
//   * 3-D source planes: float per cell, eta-major (ieta*NX*NY + ix*NY + iy),
//     holding e0 * norm_factor (the cell-independent factor of
//     convertInitialDensity, src/FreeStream.cpp:302-303,346).
//   * T^{mu nu}: 10 float planes.  Reference layout is (ix*NY + iy)*NETA +
//     ieta; the 3-D stream kernel writes an eta-major layout
//     (ieta*NX*NY + ix*NY + iy) which the Landau kernel reads directly.
//   * Outputs: 17 double planes in output_to_vectors order.
#ifndef FSM_INTERNAL_H_
#define FSM_INTERNAL_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FSM_HBARC_F 0.197326938f

// The reference's keep/drop decision for a shifted sample
// (src/FreeStream.cpp:163,187,198,204,213) with its own float operations and
// rounding: x = (float)i*D + xmin; x_new = x - shift; f = floor((x_new -
// xmin)/D); kept iff f in [1, n-2].  Contraction is off here: HIP's
// __fadd_rn/__fmul_rn are plain operators that the compiler may fuse into an
// FMA, which changes the decision when a shift is a whole number of cells.
__device__ __forceinline__ bool fsm_ref_mask_ok(int i, float shift, float mn, float h, int n) {
#pragma clang fp contract(off)
  const float x = (float)i * h + mn;
  const float xn = x - shift;
  const float f = floorf((xn - mn) / h);
  return (f >= 1.0f) && (f < (float)(n - 1));
}

// The same decision for a sample whose constant-stencil floor index is g:
// only g within one cell of the mask boundary can round either way.
__device__ __forceinline__ bool fsm_mask_ok(int i, int g, float shift, float mn, float h, int n) {
  if ((unsigned)g <= 1u || (unsigned)(g - (n - 2)) <= 1u) return fsm_ref_mask_ok(i, shift, mn, h, n);
  return (unsigned)(g - 1) <= (unsigned)(n - 3);
}

// XCD-aware work order.  An MI355X dispatches workgroups to its 8 XCDs round
// robin (workgroup b on XCD b % 8), and each XCD has its own L2.  Sending a
// contiguous range of work items to each XCD keeps neighbouring tiles, which
// read overlapping knot boxes, in one L2.  A bijection of [0, n) when every
// index b of a grid-stride loop has b % 8 equal to its workgroup's XCD (the
// grid covers n, or its size is a multiple of 8).
__device__ __forceinline__ int xcd_order(int b, int n) {
  const int q = n >> 3, rem = n & 7, x = b & 7, j = b >> 3;
  return x * q + (x < rem ? x : rem) + j;
}

// Memory-space views of pointers held in the event descriptors.  A pointer
// loaded from memory is generic, and the compiler emits flat loads through
// it; a flat load counts in both the vector-memory and the LDS/scalar wait
// counters, so any wait for an LDS read also waits for every flat load in
// flight, and prefetches issued ahead of an LDS-heavy loop are waited for at
// its first LDS read.  Global loads count only in the vector-memory counter.
// Scalar loads (constant address space) are for wave-uniform addresses of
// data no kernel writes during the launch (plan tables, descriptors).
#define FSM_AS_GLOBAL __attribute__((address_space(1)))
#define FSM_AS_CONST __attribute__((address_space(4)))
typedef unsigned fsm_v4u __attribute__((ext_vector_type(4)));
typedef unsigned fsm_v2u __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint4 fsm_gload(const uint4 *p) {
  const fsm_v4u v = *(const FSM_AS_GLOBAL fsm_v4u *)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float4 fsm_gload(const float4 *p) {
  const fsm_v4u v = *(const FSM_AS_GLOBAL fsm_v4u *)p;
  return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
}
__device__ __forceinline__ uint2 fsm_gload(const uint2 *p) {
  const fsm_v2u v = *(const FSM_AS_GLOBAL fsm_v2u *)p;
  return make_uint2(v.x, v.y);
}
__device__ __forceinline__ unsigned fsm_gload(const unsigned *p) { return *(const FSM_AS_GLOBAL unsigned *)p; }
__device__ __forceinline__ float fsm_gload(const float *p) { return *(const FSM_AS_GLOBAL float *)p; }
__device__ __forceinline__ void fsm_gstore(float *p, float v) { *(FSM_AS_GLOBAL float *)p = v; }
// Element idx of a plane whose base is wave-uniform (kept in scalar
// registers): a global store with a 32-bit per-lane byte offset, no per-lane
// 64-bit address arithmetic.  The byte offset is formed in 32 bits, so the
// plane holds fewer than 2^29 doubles (2^30 floats): a cell index, far below.
__device__ __forceinline__ void fsm_gstore_at(double *base, unsigned idx, double v) {
  *(FSM_AS_GLOBAL double *)((FSM_AS_GLOBAL char *)base + (idx << 3)) = v;
}
__device__ __forceinline__ void fsm_gstore_at(float *base, unsigned idx, float v) {
  *(FSM_AS_GLOBAL float *)((FSM_AS_GLOBAL char *)base + (idx << 2)) = v;
}
__device__ __forceinline__ void fsm_gstore(uint4 *p, uint4 v) {
  *(FSM_AS_GLOBAL fsm_v4u *)p = fsm_v4u{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ void fsm_gstore(float4 *p, float4 v) {
  *(FSM_AS_GLOBAL fsm_v4u *)p =
      fsm_v4u{__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z), __float_as_uint(v.w)};
}
// Copies 16 bytes per lane from global memory straight into LDS
// (global_load_lds_dwordx4): lane l of the wave writes lds_base[l], with
// lds_base wave-uniform and 16-byte aligned.  Inactive lanes write nothing.
// Counts in the vector-memory wait counter.
__device__ __forceinline__ void fsm_lds_dma16(const void *src, void *lds_base) {
  __builtin_amdgcn_global_load_lds((FSM_AS_GLOBAL void *)const_cast<void *>(src),
                                   (__attribute__((address_space(3))) void *)lds_base, 16, 0, 0);
}
// Scalar load of a trivially copyable descriptor (size a multiple of 4 bytes).
template <typename T>
__device__ __forceinline__ T fsm_sload(const T *p) {
  static_assert(sizeof(T) % 4 == 0, "scalar loads move whole dwords");
  unsigned w[sizeof(T) / 4];
  const FSM_AS_CONST unsigned *s = (const FSM_AS_CONST unsigned *)p;
#pragma unroll
  for (unsigned k = 0; k < sizeof(T) / 4; ++k) w[k] = s[k];
  return __builtin_bit_cast(T, w);
}

// --------------------------------------------------------------- Landau ----
#define FSM_NUM_OUTPUTS 17
// 3-D tile of the Landau kernel: FSML_TE eta planes x FSML_TX transverse
// cells (one cell per thread); outputs staged in LDS with a padded pitch.
#define FSML_TE 16
#define FSML_TX 16
#define FSML_PITCH (FSML_TE + 1)
#define FSML_PLANE (FSML_TX * FSML_PITCH)
struct EventLandau {
  int dim;
  int nxy, neta;        // for the eta-major T layout
  int t_eta_major;      // 1: T is eta-major (3-D stream output)
  int eos_type;
  int viscous;
  int fill_outputs;     // OUTPUTFORMAT is 2 or 3
  int eta_zero;         // 1: T^{tau eta}, T^{x eta}, T^{y eta}, T^{eta eta} are zero by construction (a 2-D
                        // event's tensor from the stream kernel): the 3x3 Landau frame
  float tau;
  double inv_tau2;      // 1.0 / (double)(TAU * TAU), the per-event factor of pi^{eta eta} (src/LandauMatch.cpp:365)
  const float *T;
  long t_plane;
  int t_ny, t_row;      // 2-D T with a padded row pitch: cell is = ix * t_ny + iy at ix * t_row + iy
                        // (t_row = 0: T is in the reference layout or eta-major)
  double *out;          // 17 planes of dim doubles (may be null)
  long out_plane;
  float *fe, *fu, *fp, *fPi, *fpi;  // optional float outputs
  const float *J;       // baryon current J^tau, J^x, J^y, J^eta (T's layout and plane stride), or null
  double *out_b;        // n_B, V^tau, V^x, V^y, V^eta: 5 planes of out_plane doubles (with J)
};

// ---------------------------------------------------------------- 2-D ----
// Cells per lane along x in the 2-D stream kernel; one wave covers 64
// consecutive iy.  Tile = (4 waves * FSM2D_M) x 64 cells.
#define FSM2D_M 4
#define FSM2D_WAVES 4
#define FSM2D_TX (FSM2D_WAVES * FSM2D_M)
#define FSM2D_TY 64
#define FSM2D_THREADS (FSM2D_WAVES * 64)

// Band half-width of the natural-spline derivative operator.  The operator
// is the inverse of the tridiagonal (1,4,1) spline system times a central
// difference; its entries decay as (2 - sqrt 3)^|i-m| = 0.268^|i-m|, so
// entries beyond 16 are below 1e-9 of the largest (checked on the host when
// the tables are built) -- far below float resolution.
#define FSM_SPLINE_BAND 16
#define FSM_SPLINE_W (2 * FSM_SPLINE_BAND + 1)
#define FSMSP_TX 16
#define FSMSP_TY 32
// Batched spline tiles: FSMSP2_TX rows x FSMSP_TY columns (several flag tiles).
#define FSMSP2_TX 64

// One azimuthal sample of the 2-D stencil.  For a fixed phi_j the source
// point x - DTAU cos(phi_j) is the same shift for every cell
// (src/FreeStream.cpp:184-189), so the floor offset and the interpolation
// weights are per-j constants.
struct Stencil2D {
  int ox, oy;       // knot offset of the lower-left corner, in cells
  float hx[4];      // Hermite weights in x: value0, slope0, value1, slope1
  float hy[4];      // Hermite weights in y
  float tx, ty;     // bilinear fractions
  float c, s;       // cos(phi_j), sin(phi_j) as the reference computes them
  float cc, cs, ss; // hypertrig products c*c, c*s, s*s
  float dxj, dyj;   // cos(phi_j)*DTAU, sin(phi_j)*DTAU (float, as the reference)
  float pad;
};
static_assert(sizeof(Stencil2D) == 80, "Stencil2D is staged in LDS as 5 float4");

// A contiguous range of azimuthal samples whose stencil offsets share a
// bounding box; the 2-D kernel stages that box of knots in LDS once per range.
struct StencilGroup {
  int j0, j1;  // [j0, j1)
  int bx0, bx1, by0, by1;  // knot offset box: ox in [bx0, bx1-1], knots up to bx1
  int pad0, pad1;
};

// Arguments of k_stage_cells: one of the reference's per-cell stages on
// planes of `dim` floats in the reference layout.
enum { FSM_STAGE_PRESSURE = 0, FSM_STAGE_BULK = 1, FSM_STAGE_SHEAR = 2 };
struct StageCells {
  long dim;
  int stage;
  int eos_type;
  float tau;
  double inv_tau2;      // 1.0 / (double)(TAU * TAU), TAU * TAU a float product
  const float *T;       // 10 planes (bulk, shear)
  const float *e;       // energy density
  const float *u;       // 4 planes (shear)
  const float *p;       // pressure (bulk, shear)
  const float *Pi;      // bulk pressure (shear)
  float *p_out, *Pi_out, *pi_out;  // pi_out: 10 planes
};

struct Event2D {
  int nx, ny;
  int ngroups;
  int lds_floats4;      // LDS float4 elements needed by the largest group
  float out_scale;      // DIM_RAP (identical rapidity copies in 2-D)
  float d_phip;         // (float)(2 pi / DIM_PHIP)
  float tau;            // TAU
  int tiles_x;          // tiles along x
  int tiles_y;          // tiles along y
  int variant;          // stream kernel shape (fsm_stream2d.hip: 0 = M4/JS1, 1 = M4/JS4, 2 = M2/JS4)
  float xmin, ymin;     // grid origin as getX/getY compute it
  float dx, dy;
  const float *e_gev;   // input, GeV/fm^3
  float *zk;            // batches: knot values z of e_gev (written by the knot pass, read by the batch
                        // spline kernel); the one-event kernels convert e_gev themselves
  float4 *planes;       // spline planes (written by the spline kernel)
  const float *gx;      // [nx][FSM_SPLINE_W] derivative band weights along x
  const float *gy;      // [ny][FSM_SPLINE_W]
  int cxlo, cxhi;       // rows [cxlo, cxhi] of gx are bitwise the same weights (away from the edges)
  int cylo, cyhi;       // columns [cylo, cyhi] of gy likewise
  const Stencil2D *st;  // [DIM_PHIP]
  const StencilGroup *groups;
  float *T;             // 10 planes: cell (ix, iy) at ix * t_row + iy
  long t_plane;         // plane stride of T (>= nx * t_row)
  int t_row;            // row pitch of T in cells: ny (the reference layout), or ny rounded up to 32
                        // (a batch's internal T: every tile's 32-row run starts on a 128-byte line)
  // ---- mixed-precision quadrature (MFMA tiles) ----
  int mfma;             // 1: MFMA tile kernel; spline kernel also writes planes_h and sp_flags
  int n_moff;           // knot offsets carrying at least one MFMA sample
  int n_mgroups;        // offset groups (StencilGroup with j0/j1 = offset range)
  int n_amb;            // samples whose floor index is rounding-dependent (edge-mask correction)
  int mfma_lds4;        // 16-byte LDS elements of the largest offset group (box + weights + records)
  int rbx0, rbx1, rby0, rby1;  // knot box of all offsets (as StencilGroup.bx0..by1)
  int sp_tiles_y;       // spline tiles along y (flag index = sx * sp_tiles_y + sy)
  uint4 *planes_h;      // per knot 8 halves: scale*{z, zy, zx, zxy} hi, then lo
  unsigned *sp_flags;   // per spline tile: nonzero if a knot is unsafe for the MFMA form
  float *mscale;        // per event: FSM_KS_PARTS partial max |z|, then the power-of-two knot scale
  const int4 *moffs;    // [n_moff] (ox, oy, LDS delta of the offset in its group's knot box, -)
  const uint4 *bmat;    // per MFMA FSM_BMAT_U4 uint4: the B operand packed, 4 halves per lane
  const StencilGroup *mgroups;
  const int *amb;       // [n_amb] indices into st
  // ---- charge pass of the baryon extension ----
  int charge;           // 1: the input is n_B (fm^-3) and z = n_B * in_norm (convertInitialChargeDensity)
  float in_norm;        // its normalisation (src/FreeStream.cpp:365-366)
  int raw_e;            // 1: the energy input is already the fm^-4 density (IC_ENERGY != 5)
  int out_j;            // 1: write J^tau, J^x, J^y, J^eta (4 planes of T) instead of T^{mu nu}
  float jscale;         // DRAP of calculateBaryonCurrent (src/LandauMatch.cpp:140)
  int skip_eta_planes;  // 1: T is workspace read only by the 3x3 Landau frame, which does not read
                        // T^{tau eta}, T^{x eta}, T^{y eta}, T^{eta eta} (zeros): they are not written
  // ---- fused Landau stage (events without a baryon current) ----
  int fuse_landau;      // 1: the tile epilogue runs the 3x3 Landau frame on the moments in registers
                        // and writes the 17 output planes (l_out); T is not written
  int l_eos, l_viscous, l_fill;  // EventLandau::eos_type, viscous, fill_outputs
  double l_inv_tau2;    // EventLandau::inv_tau2
  double *l_out;        // EventLandau::out (17 planes of l_out_plane doubles)
  long l_out_plane;
};

// The MFMA form of the 2-D quadrature.  For all samples j sharing a knot
// offset o the interpolation and the angular moments are linear in the 16
// knot values of the patch, so
//   T_m(cell) = sum_o sum_{corner,comp} K[cell+o+corner][comp] * W_o[corner,comp][m],
//   W_o = sum_{j in o} hermite_j[corner,comp] * P_j[m],
// one 16x16x32 FP16 MFMA per 16 cells and offset, with the knots and the
// weights split into hi + lo halves ((K_h + K_l)(W_h + W_l), all four products).  Knots are
// multiplied by a per-event power of two before the split, chosen from the
// event's largest |z| so that the largest knot value is near
// FSM_MFMA_ZTARGET (the fp16 subnormal floor then sits ~2^-36 below it; the
// scale is undone in FP32 at the end).  A spline tile is flagged when a
// scaled component exceeds FSM_MFMA_MAXABS, or when a patch could go <= 0
// (Bernstein control points, where the reference falls back to bilinear),
// and MFMA tiles that read a flagged knot run the FP32 sample loop instead.
#define FSM_MFMA_ZTARGET 4096.0f
// 16-byte units of one packed MFMA B operand (64 lanes x 8 bytes)
#define FSM_BMAT_U4 32
#define FSM_KS_PARTS 64
#define FSM_MFMA_MAXABS 30000.0f
#define FSM_MFMA_POS_EPS 1.0e-5f

// ---------------------------------------------------------------- 3-D ----
#define FSM3D_THREADS 256
#define FSM3D_M 4
#define FSM3D_TILE (FSM3D_THREADS * FSM3D_M)
// LDS budget of the staged source rows of one 3-D tile (floats); the window
// grows past it only when a tile's own rows need more.
#define FSM3D_WIN_FLOATS 6144

// Per (ieta, irap) quantities of the 3-D path.  The source rapidity plane,
// the eta interpolation fraction and the transverse shift magnitude D depend
// on (ieta, irap) only (src/FreeStream.cpp:174-195).
struct RapSlot {
  int valid;        // eta-mask passes and D is finite
  int eta0;         // floor of the source eta index
  float feta;       // eta fraction
  float D;          // TAU cosh(rap-eta_new) - TAU0 cosh(rap-eta)
  float rf0, rf1;   // rapidity factor of convertInitialDensity at planes eta0, eta0+1
  float jac;        // rapidity Jacobian of calculateStressTensor
  float inv_ch;     // 1/cosh(rap-eta)
  float inv_ch2;    // 1/cosh^2(rap-eta)
  float th_t;       // tanh(rap-eta)/TAU
  float th_tch;     // tanh(rap-eta)/(TAU cosh(rap-eta))
  float th2_t2;     // tanh^2(rap-eta)/TAU^2
  int s0, ns;       // this slot's live azimuth samples: Samp3D [s0, s0+ns)
  int pad0, pad1;
};

// One azimuth sample of a 3-D rapidity slot that keeps at least one cell.
// The shift (cos phi_j, sin phi_j) * D is the same for every cell, so the
// floor offset and the bilinear weights are per-sample constants.  The
// reference keeps a sample iff the floor of its float grid index lies in
// [1, N-2] (src/FreeStream.cpp:213); that floor is monotone in the cell
// index, so the kept cells form one interval per axis, found on the host with
// the reference's own float expressions.  The kernel reads the fields it
// needs per sample already combined: the kept box packed for a 16-bit-lane
// test against (ix << 16 | iy), and the LDS byte offset of the sample's
// source corner relative to the cell in the staged-row window.
struct Samp3D {
  unsigned mlo;              // (xlo << 16) | ylo: first kept cell along each axis
  unsigned mext;             // ((xhi - xlo) << 16) | (yhi - ylo): kept extent - 1 along each axis
  int d0;                    // 4 * (ox * DIM_Y + oy): byte offset in a window of pitch DIM_Y
  float w00, w01, w10, w11;  // bilinear weights of Q[ix+ox][iy+oy], [..][iy+oy+1], [ix+ox+1][..], [..+1][..+1]
  float c, s, cc, cs;        // cos phi_j, sin phi_j, c*c, c*s
  int ox, oy;                // floor offset (cells) of the shifted point
  int xlo, xhi;              // kept cells along x: ix in [xlo, xhi] (host window planning)
  int pad;
};
static_assert(sizeof(Samp3D) == 64, "Samp3D is read with one scalar load");

struct Event3D {
  int nx, ny, neta, nrap, nphi;
  int tiles;            // transverse tiles per eta plane
  int lds_floats;       // LDS floats of the staged-row window and its zero patch ((win_rows + 2) * NY)
  int win_rows;         // rows of the window
  float d_phip;
  float inv_dx, inv_dy; // 1/DX, 1/DY: shift in cells = -(cos(phi) * D) * inv_dx
  float dx, dy;
  float xmin, ymin;
  const float *e_gev;   // input GeV/fm^3, reference layout
  float *src;           // eta-major e0*norm planes
  const RapSlot *slots; // [neta][nrap]
  const Samp3D *samp;   // live samples of every slot, each slot's sorted by (ox, oy)
  const int2 *srange;   // [neta*nrap][tiles]: the slot's staged-row windows [x, y) for the tile
  const int4 *win;      // windows: samples [x, y) of the slot (relative to s0), served by window rows
                        // [z, z + w) of the eta-blended source plane
  const int *eta_order; // eta planes by decreasing work (the one-event launch's order)
  float *T;             // 10 planes, eta-major layout
  long t_plane;
  float norm_factor;    // energy: convertInitialDensity norm; charge pass: convertInitialChargeDensity norm
  int charge;           // 1: the input is n_B: src = n_B * norm_factor (no hbarc, no floor)
  int raw_e;            // 1: the energy input is already the fm^-4 density (IC_ENERGY != 5)
};

#endif  // FSM_INTERNAL_H_
