// fsm_landau.h -- per-cell Landau matching (solveEigenSystem), EoS, bulk and
// shear parts and the output pack, as device functions shared by the Landau
// kernels (fsm_kernels.hip) and the 2-D stream kernels' fused epilogue
// (fsm_stream2d.hip).  Included inside each file's anonymous namespace.
#ifndef FSM_LANDAU_H_
#define FSM_LANDAU_H_

// ======================================================================
// Landau matching + EoS + viscous parts + output pack
// ======================================================================
// Landau frame of one cell by Jacobi sweeps.  In the orthonormal frame (eta
// components scaled by TAU) the matrix T^mu_nu = T^{mu rho} g_{rho nu} of
// src/LandauMatch.cpp:177-214 becomes A eta with A symmetric; its timelike
// eigenvector is the frame in which A^{0i} = 0.  Each step is the exact
// Lorentz boost in the (0,i) plane that zeroes A^{0i}:
//   tanh(2 theta) = -2 A^{0i} / (A^{00} + A^{ii}),
// i.e. the minimiser of u.A.u along that boost, which is the closed form
// e = (a-c)/2 + sqrt((a+c)^2/4 - b^2) of the 2x2 block.  Cyclic sweeps over
// i = x, y, eta converge to the Landau frame; the product of the inverse
// boosts maps (1,0,0,0) to u.  Selection rules of the reference
// (src/LandauMatch.cpp:228-291): e > 0, u timelike, u^tau < 100, else
// e = 1e-3 and u = (1,0,0,0).
// Returns 1 when a Landau frame is accepted (e, u written), 0 when the
// reference's fallback applies (e = 1e-3, u = (1,0,0,0) written).
// tol2: squared relative off-diagonal residual of convergence; loose2: the
// residual still accepted after the last sweep.
// Reciprocal square root in FP64: the hardware estimate refined by two
// Newton steps (each doubles the bits).  Arguments are positive and finite.
__device__ __forceinline__ double jrsq(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double r = fma(-(x * y), y, 1.0);
  y = fma(0.5 * y, r, y);
  r = fma(-(x * y), y, 1.0);
  return fma(0.5 * y, r, y);
}

// These sweeps decide the cells where the characteristic quartic (below)
// does not settle or fails a rule.  (2-D events, whose tensors from the
// stream kernel have zero eta components, take the closed 3x3 form below:
// landau_cubic.)
template <typename R>
__device__ int landau_jacobi(const float t[10], float TAU, R tol2, R loose2, float *e_out, float u_out[4]) {
  constexpr int N = 4;
  const R tau = (R)TAU;
  R A[N][N], M[N][N];
  A[0][0] = t[0];
  A[0][1] = A[1][0] = t[1];
  A[0][2] = A[2][0] = t[2];
  A[1][1] = t[4];
  A[1][2] = A[2][1] = t[5];
  A[2][2] = t[7];
  A[0][3] = A[3][0] = tau * (R)t[3];
  A[1][3] = A[3][1] = tau * (R)t[6];
  A[2][3] = A[3][2] = tau * (R)t[8];
  A[3][3] = tau * tau * (R)t[9];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) M[i][j] = (i == j) ? (R)1 : (R)0;
  auto off_diag = [&]() {
    return A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[0][3] * A[0][3];
  };
  bool ok = true, conv = false;
  for (int sweep = 0; sweep < 60 && ok; ++sweep) {
    const R off = off_diag();
    if (off <= tol2 * A[0][0] * A[0][0]) {
      conv = true;
      break;
    }
#pragma unroll
    for (int i = 1; i < N; ++i) {
      const R b = A[0][i];
      if (b == (R)0) continue;
      const R s = A[0][0] + A[i][i], d = A[0][0] - A[i][i];
      const R disc = (s - (R)2 * fabs(b)) * (s + (R)2 * fabs(b));
      if (!(s > (R)0) || !(disc > (R)0)) {
        ok = false;
        break;
      }
      const R rq = jrsq(disc);
      const R sq = disc * rq;        // sqrt(disc)
      const R ch2 = s * rq;          // cosh 2theta
      const R sh2 = (R)-2 * b * rq;  // sinh 2theta
      const R x = (R)0.5 * (ch2 + (R)1);
      const R rx = jrsq(x);
      const R ch = x * rx;                // cosh theta = sqrt(x)
      const R sh = (R)0.5 * sh2 * rx;     // sinh theta = sinh 2theta / (2 cosh theta)
      // the boosted 2x2 block in closed form: diag((d + sq)/2, (sq - d)/2)
      A[0][0] = (R)0.5 * (d + sq);
      A[i][i] = (R)0.5 * (sq - d);
      A[0][i] = A[i][0] = (R)0;
#pragma unroll
      for (int k = 1; k < N; ++k) {
        if (k == i) continue;
        const R a0 = A[0][k], ai = A[i][k];
        A[0][k] = A[k][0] = ch * a0 + sh * ai;
        A[i][k] = A[k][i] = sh * a0 + ch * ai;
      }
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const R m0 = M[k][0], mi = M[k][i];
        M[k][0] = ch * m0 - sh * mi;
        M[k][i] = ch * mi - sh * m0;
      }
    }
  }
  if (ok && !conv) {
    const R off = off_diag();
    conv = off <= loose2 * A[0][0] * A[0][0];
  }
  const R e = A[0][0];
  const R u0 = M[0][0];
  if (ok && conv && e > (R)0 && u0 < (R)100 && isfinite(u0)) {
    *e_out = (float)e;
    u_out[0] = (float)u0;
    u_out[1] = (float)M[1][0];
    u_out[2] = (float)M[2][0];
    u_out[3] = (float)(M[3][0] / tau);
    return 1;
  }
  *e_out = 1.0e-3f;
  u_out[0] = 1.0f;
  u_out[1] = u_out[2] = u_out[3] = 0.0f;
  return 0;
}

// FP64 solve: converged when the off-diagonal row is below 1e-12 of A^{00}
// (on sampled cells of a lumpy 2-D and a 3-D event its float results equal
// those of a 1e-14 criterion; the eigenvalue error is quadratic in the
// residual).
__device__ void landau_fp64(const float t[10], float TAU, float *e_out, float u_out[4]) {
  (void)landau_jacobi<double>(t, TAU, 1.0e-24, 1.0e-12, e_out, u_out);
}

__device__ __forceinline__ double lrcp(double x);  // below, with landau_cubic

// Landau frame of a 4x4 cell from the characteristic quartic (the eigenpair
// that solveEigenSystem, src/LandauMatch.cpp:145-331, takes from a general
// eigen-solver).  In the orthonormal frame (eta components times tau) S = T
// is symmetric, g = diag(1, -1, -1, -1), and T g u = e u is (S - e g) w = 0
// with w = g u.  p(l) = det(S - l g) = -l^4 + c3 l^3 + c2 l^2 + c1 l + c0,
// whose coefficients are signed sums of the principal minors of S.  For a
// tensor of free-streamed massless particles (a positive sum of null dyads)
// the roots are e > 0 and -P_i <= 0, so e is the largest root, and
// T^{tau tau} and T(n, n) at the velocity estimate n bound it from above;
// Halley's iteration from there converges to it (cubically near the root).
// w is any row of adj(S - e g) = c w w^T; row 0 is the longest (|w^0| =
// u^tau > |u^i|) and needs only the 2x2 minors of rows 2 and 3.  S is first
// scaled by a power of two near 1 / T^{tau tau} (exact), so no product of
// up to six components leaves the float range.
// FP64 (an FP64 FMA issues at the FP32 rate on this GPU).
// Returns 1 when the frame is found and passes the reference's rules
// (e > 0, u timelike, u^tau < 100); 0 when the iteration did not settle or a
// rule fails, and the caller then decides the cell with the Jacobi sweeps.
__device__ int landau_quartic(const float t[10], float TAU, float *e_out, float u_out[4]) {
  using R = double;
  const R tau = (R)TAU;
  const R t00 = (R)t[0];
  if (!(t00 > (R)0) || !isfinite(t00)) return 0;
  int ex;
  (void)frexp(t00, &ex);
  const R sc = ldexp((R)1, -ex);  // t00 * sc in [0.5, 1)
  const R s00 = t00 * sc, s01 = (R)t[1] * sc, s02 = (R)t[2] * sc, s03 = (tau * (R)t[3]) * sc;
  const R s11 = (R)t[4] * sc, s12 = (R)t[5] * sc, s13 = (tau * (R)t[6]) * sc;
  const R s22 = (R)t[7] * sc, s23 = (tau * (R)t[8]) * sc;
  const R s33 = (tau * tau * (R)t[9]) * sc;
  // principal minors
  const R q01 = s01 * s01, q02 = s02 * s02, q03 = s03 * s03, q12 = s12 * s12, q13 = s13 * s13, q23 = s23 * s23;
  const R m01 = s00 * s11 - q01, m02 = s00 * s22 - q02, m03 = s00 * s33 - q03;
  const R m12 = s11 * s22 - q12, m13 = s11 * s33 - q13, m23 = s22 * s33 - q23;
  const R m123 = s11 * m23 - s22 * q13 - s33 * q12 + (R)2 * s12 * s23 * s13;
  const R m023 = s00 * m23 - s22 * q03 - s33 * q02 + (R)2 * s02 * s23 * s03;
  const R m013 = s00 * m13 - s11 * q03 - s33 * q01 + (R)2 * s01 * s13 * s03;
  const R m012 = s00 * m12 - s11 * q02 - s22 * q01 + (R)2 * s01 * s12 * s02;
  // det S from the 2x2 minors of rows 0, 1 and of rows 2, 3
  const R a01 = s00 * s11 - s01 * s01, a02 = s00 * s12 - s02 * s01, a03 = s00 * s13 - s03 * s01;
  const R a12 = s01 * s12 - s02 * s11, a13 = s01 * s13 - s03 * s11, a23 = s02 * s13 - s03 * s12;
  const R b01 = s02 * s13 - s12 * s03, b02 = s02 * s23 - s22 * s03, b03 = s02 * s33 - s23 * s03;
  const R b12 = s12 * s23 - s22 * s13, b13 = s12 * s33 - s23 * s13, b23 = s22 * s33 - s23 * s23;
  const R c0 = a01 * b23 - a02 * b13 + a03 * b12 + a12 * b03 - a13 * b02 + a23 * b01;
  const R c1 = m012 + m013 + m023 - m123;
  const R c2 = m01 + m02 + m03 - m12 - m13 - m23;
  const R c3 = s00 - s11 - s22 - s33;
  // start: T(n, n) at n = gamma (1, v), v = T^{0i} / (T^{00} + tr/3)
  const R rd = lrcp(s00 + (s11 + s22 + s33) * (R)(1.0 / 3.0));
  const R vx = s01 * rd, vy = s02 * rd, vz = s03 * rd;
  const R g2 = lrcp((R)1 - vx * vx - vy * vy - vz * vz);
  const R tnn = g2 * (s00 - (R)2 * (vx * s01 + vy * s02 + vz * s03) + vx * vx * s11 + vy * vy * s22 +
                      vz * vz * s33 + (R)2 * (vx * vy * s12 + vx * vz * s13 + vy * vz * s23));
  R lam = fmin(s00, tnn);  // fmin: T^{00} where the estimate is not a number
  // After a step below tol the remaining error is of order tol^3: rounding.
  const R tol = (R)1.0e-10;
  bool conv = false;
  for (int it = 0; it < 16; ++it) {
    const R p = (((c3 - lam) * lam + c2) * lam + c1) * lam + c0;
    const R dp = (((R)3 * c3 - (R)4 * lam) * lam + (R)2 * c2) * lam + c1;
    const R d2 = ((R)6 * c3 - (R)12 * lam) * lam + (R)2 * c2;
    const R st = (R)2 * p * dp * lrcp((R)2 * dp * dp - p * d2);
    lam -= st;
    if (!(fabs(st) > tol * fabs(lam))) {
      conv = fabs(st) <= tol * fabs(lam);  // false for a NaN
      break;
    }
  }
  if (!conv || !(lam > (R)0)) return 0;
  // row 0 of adj(S - lam g): rows 2, 3 of B = S - lam g and their 2x2 minors
  const R b11 = s11 + lam, b22 = s22 + lam, b33 = s33 + lam;
  const R n01 = s02 * s13 - s12 * s03, n02 = s02 * s23 - b22 * s03, n03 = s02 * b33 - s23 * s03;
  const R n12 = s12 * s23 - b22 * s13, n13 = s12 * b33 - s23 * s13, n23 = b22 * b33 - s23 * s23;
  const R w0 = b11 * n23 - s12 * n13 + s13 * n12;
  const R w1 = s01 * n23 - s12 * n03 + s13 * n02;    // -C01 = u^x direction
  const R w2 = -(s01 * n13 - b11 * n03 + s13 * n01);  // -C02
  const R w3 = s01 * n12 - b11 * n02 + s12 * n01;    // -C03
  const R ml = w0 * w0 - (w1 * w1 + w2 * w2 + w3 * w3);
  if (!(ml > (R)0) || !isfinite(ml)) return 0;
  R f = jrsq(ml);
  if (w0 < (R)0) f = -f;
  const R u0 = w0 * f;
  if (!(u0 < (R)100)) return 0;
  *e_out = (float)ldexp(lam, ex);
  u_out[0] = (float)u0;
  u_out[1] = (float)(w1 * f);
  u_out[2] = (float)(w2 * f);
  u_out[3] = (float)(w3 * f / tau);
  return 1;
}

// Landau frame of a 4x4 cell, both paths: the quartic, and the Jacobi sweeps
// where the quartic does not settle or a rule fails (those decide the
// reference's fallback).
__device__ void landau_frame4(const float t[10], float TAU, float *e_out, float u_out[4]) {
  if (!landau_quartic(t, TAU, e_out, u_out)) landau_fp64(t, TAU, e_out, u_out);
}

// Landau frame of a 2-D event's tensor (eta components zero), FP64.  With
// A = T eta (A^mu_nu = T^{mu alpha} eta_{alpha nu}, 3x3), e is an eigenvalue
// of A with a timelike eigenvector, and for a tensor of free-streamed
// massless particles (a positive sum of null dyads) it is the largest root of
// the characteristic cubic det(A - lambda) = 0: the others are -P_1, -P_2 with
// pressures P_i >= 0.  e is also the minimum of T(n, n) over unit timelike n,
// so T(n, n) at n = gamma (1, v), v = T^{0i} / (T^{00} + (T^{xx} + T^{yy})/3),
// lies above every root; Halley's iteration from there converges to e (five
// or six steps on the lumpy and noisy test events, from the first; cubic
// convergence near the root).  u is the null vector of A - e: the largest of
// the three cross products of its rows, normalised to u.u = 1 with u^tau > 0.
// The reference's rules (src/LandauMatch.cpp:222-291) then decide: e > 0,
// u timelike, u^tau < 100, else e = 1e-3, u = (1, 0, 0, 0).  Returns 1 when
// the frame is found.
__device__ __forceinline__ double lrcp(double x) {
  // hardware reciprocal estimate refined by two Newton steps
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  return fma(fma(-x, r, 1.0), r, r);
}
__device__ int landau_cubic(const float t[10], float *e_out, float u_out[4]) {
  const double t0 = t[0], t1 = t[1], t2 = t[2], t4 = t[4], t5 = t[5], t7 = t[7];
  // A = T eta: rows (t0, -t1, -t2), (t1, -t4, -t5), (t2, -t5, -t7)
  const double a00 = t0, a01 = -t1, a02 = -t2, a10 = t1, a11 = -t4, a12 = -t5, a20 = t2, a21 = -t5, a22 = -t7;
  const double tr = a00 + a11 + a22;
  const double m12 = a11 * a22 - a12 * a21;
  const double c2 = (a00 * a11 - a01 * a10) + (a00 * a22 - a02 * a20) + m12;
  const double det = a00 * m12 - a01 * (a10 * a22 - a12 * a20) + a02 * (a10 * a21 - a11 * a20);
  const double rd = lrcp(t0 + (t4 + t7) * (1.0 / 3.0));
  const double vx = t1 * rd, vy = t2 * rd;
  const double g2 = lrcp(1.0 - vx * vx - vy * vy);
  const double tnn = g2 * (t0 - 2.0 * (vx * t1 + vy * t2) + vx * vx * t4 + 2.0 * vx * vy * t5 + vy * vy * t7);
  double lam = fmin(t0, tnn);  // fmin: T^{00} where the estimate is not a number
  for (int it = 0; it < 12; ++it) {
    const double p = ((lam - tr) * lam + c2) * lam - det;
    const double dp = (3.0 * lam - 2.0 * tr) * lam + c2;
    const double d2 = 6.0 * lam - 2.0 * tr;
    const double st = 2.0 * p * dp * lrcp(2.0 * dp * dp - p * d2);
    lam -= st;
    if (!(fabs(st) > 1.0e-13 * fabs(lam))) break;  // converged (the next step is below rounding), or not a number
  }
  // rows of A - e and their cross products
  const double b00 = a00 - lam, b11 = a11 - lam, b22 = a22 - lam;
  const double p0 = a01 * a12 - a02 * b11, p1 = a02 * a10 - b00 * a12, p2 = b00 * b11 - a01 * a10;  // r0 x r1
  const double q0 = a01 * b22 - a02 * a21, q1 = a02 * a20 - b00 * b22, q2 = b00 * a21 - a01 * a20;  // r0 x r2
  const double s0 = b11 * b22 - a12 * a21, s1 = a12 * a20 - a10 * b22, s2 = a10 * a21 - b11 * a20;  // r1 x r2
  const double np = p0 * p0 + p1 * p1 + p2 * p2, nq = q0 * q0 + q1 * q1 + q2 * q2, ns = s0 * s0 + s1 * s1 + s2 * s2;
  const bool use_p = np >= nq && np >= ns, use_q = !use_p && nq >= ns;
  const double v0 = use_p ? p0 : use_q ? q0 : s0, v1 = use_p ? p1 : use_q ? q1 : s1, v2 = use_p ? p2 : use_q ? q2 : s2;
  const double ml = v0 * v0 - (v1 * v1 + v2 * v2);
  // the reference's fallback unless the frame passes its rules; every output
  // is written once, after the branch (no conditionally written array)
  float eo = 1.0e-3f, w0 = 1.0f, w1 = 0.0f, w2 = 0.0f;
  int found = 0;
  if (lam > 0.0 && ml > 0.0 && isfinite(lam) && isfinite(ml)) {
    double f = jrsq(ml);
    if (v0 < 0.0) f = -f;
    const double u0 = v0 * f;
    if (u0 < 100.0) {
      eo = (float)lam;
      w0 = (float)u0;
      w1 = (float)(v1 * f);
      w2 = (float)(v2 * f);
      found = 1;
    }
  }
  *e_out = eo;
  u_out[0] = w0;
  u_out[1] = w1;
  u_out[2] = w2;
  u_out[3] = 0.0f;
  return found;
}

__device__ float eos_pressure(float e, int eos) {
  // (float)((double)e / 3.0) in the reference; e / 3 has a period-2 binary
  // expansion (or is exact), so rounding it to double first can never land on
  // a float rounding boundary: the IEEE float quotient is the same number
  if (eos == 1) return e / 3.0f;
  if (eos == 2) {
    const float e1 = e, e2 = e * e, e3 = e2 * e, e4 = e3 * e, e5 = e4 * e, e6 = e5 * e, e7 = e6 * e,
                e8 = e7 * e, e9 = e8 * e, e10 = e9 * e, e11 = e10 * e, e12 = e11 * e;
    const float a =
        fmaf(1.9531729608963267e-11f, e12,
             fmaf(3.1188455176941583e-7f, e11,
                  fmaf(0.0009417586777847889f, e10,
                       fmaf(0.7158279081255019f, e9,
                            fmaf(141.5073484468774f, e8,
                                 fmaf(6340.448389300905f, e7,
                                      fmaf(41913.439282708554f, e6,
                                           fmaf(334334.4309240126f, e5,
                                                fmaf(1.6357487344679043e6f, e4,
                                                     fmaf(3.1729694865420084e6f, e3,
                                                          fmaf(1.077580993288114e6f, e2,
                                                               fmaf(9737.845799644809f, e1,
                                                                    -0.25181736420168666f))))))))))));
    const float b =
        fmaf(3.2581066229887368e-18f, e12,
             fmaf(5.928138360995685e-11f, e11,
                  fmaf(9.601103399348206e-7f, e10,
                       fmaf(0.002962497695527404f, e9,
                            fmaf(2.3405487982094204f, e8,
                                 fmaf(499.04919730607065f, e7,
                                      fmaf(26452.34905933697f, e6,
                                           fmaf(278581.2989342773f, e5,
                                                fmaf(1.7851642641834426e6f, e4,
                                                     fmaf(1.3512402226067686e7f, e3,
                                                          fmaf(2.0931169138134286e7f, e2,
                                                               fmaf(4.0574329080826794e6f, e1,
                                                                    45829.44617893836f))))))))))));
    return a / b;
  }
  return 0.0f;
}

// calculateBulkPressure (src/LandauMatch.cpp:332-344), one cell, in the
// reference's float expressions (contraction off).
__device__ __forceinline__ float bulk_pressure(const float t[10], float TAU, float e, float p) {
#pragma clang fp contract(off)
  const float a = t[0] - t[4] - t[7] - TAU * TAU * t[9];
  return (float)((-1.0 / 3.0) * (double)(a - e) - (double)p);
}

// calculateShearViscTensor (src/LandauMatch.cpp:345-367), one cell;
// inv_tau2 = 1.0 / (double)(TAU * TAU) with TAU * TAU a float product.
__device__ __forceinline__ void shear_tensor(const float t[10], double inv_tau2, float e, const float u[4], float p,
                                             float Pi, float pi[10]) {
#pragma clang fp contract(off)
  const float b = e + p + Pi;
  const float c = p + Pi;
  pi[0] = t[0] - u[0] * u[0] * b + c;
  pi[1] = t[1] - u[0] * u[1] * b;
  pi[2] = t[2] - u[0] * u[2] * b;
  pi[3] = t[3] - u[0] * u[3] * b;
  pi[4] = t[4] - u[1] * u[1] * b - c;
  pi[5] = t[5] - u[1] * u[2] * b;
  pi[6] = t[6] - u[1] * u[3] * b;
  pi[7] = t[7] - u[2] * u[2] * b - c;
  pi[8] = t[8] - u[2] * u[3] * b;
  pi[9] = (float)((double)(t[9] - u[3] * u[3] * b) - (double)c * inv_tau2);
}

// Results of one cell in the reference's float precision.
struct CellOut {
  float e, p, u[4], pi[10], Pi;
  float nb, V[4];  // baryon extension (with ev.J)
};

// Landau frame, EoS, bulk and shear parts of the cell with components t
// (J, with a baryon current, at index tidx of ev.J's planes); landau_cell_compute
// reads t at index tidx of ev.T's planes.  N = 3 (2-D events, eta components zero):
// the 3x3 frame from the characteristic cubic; N = 4: the 4x4 frame from the
// quartic.  Both in FP64 on both paths.
template <int N>
__device__ __forceinline__ void landau_cell_from_t(const EventLandau &ev, const float (&t)[10], long tidx,
                                                   CellOut &r) {
#pragma clang fp contract(off)
  float e, u[4];
  if constexpr (N == 3)
    landau_cubic(t, &e, u);
  else
    landau_frame4(t, ev.tau, &e, u);
  const float TAU = ev.tau;
  const float p = eos_pressure(e, ev.eos_type);
  float Pi = 0.0f;
#pragma unroll
  for (int c = 0; c < 10; ++c) r.pi[c] = 0.0f;
  if (ev.viscous) {
    Pi = bulk_pressure(t, TAU, e, p);
    shear_tensor(t, ev.inv_tau2, e, u, p, Pi, r.pi);
  }
  r.e = e;
  r.p = p;
  r.Pi = Pi;
#pragma unroll
  for (int c = 0; c < 4; ++c) r.u[c] = u[c];
  if (ev.J) {
    // src/LandauMatch.cpp:370-392: n_B = u_mu J^mu, V^mu = J^mu - n_B u^mu
    float j[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) j[c] = ev.J[c * ev.t_plane + tidx];
    const float nb = (u[0] * j[0]) - (u[1] * j[1]) - (u[2] * j[2]) - (TAU * TAU * u[3] * j[3]);
    r.nb = nb;
#pragma unroll
    for (int c = 0; c < 4; ++c) r.V[c] = j[c] - (nb * u[c]);
  }
}
template <int N>
__device__ void landau_cell_compute(const EventLandau &ev, bool live, long tidx, CellOut &r) {
  if (!live) return;
  float t[10];
#pragma unroll
  for (int c = 0; c < 10; ++c) t[c] = (N == 4 || (c != 3 && c != 6 && c != 8 && c != 9)) ? ev.T[c * ev.t_plane + tidx] : 0.0f;
  landau_cell_from_t<N>(ev, t, tidx, r);
}

// Output plane c (0..21): the float value and whether the reference scales it
// by hbarc when it fills its double vectors (src/FreestreamMilne.cpp:695-718):
// the 17 vectors in output_to_vectors order, then n_B, V^mu.
__device__ __forceinline__ float cell_value(const CellOut &r, int c) {
  if (c == 0) return r.e;
  if (c == 1) return r.p;
  if (c < 6) return r.u[c - 2];
  if (c < 16) return r.pi[c - 6];
  if (c == 16) return r.Pi;
  if (c == FSM_NUM_OUTPUTS) return r.nb;
  return r.V[c - FSM_NUM_OUTPUTS - 1];
}
__device__ __forceinline__ bool plane_scaled(int c) { return c < 2 || (c >= 6 && c < FSM_NUM_OUTPUTS); }
// (double)x * hbarc is exact in double (two 24-bit significands), so a value
// staged as float and converted later gives the same double.
__device__ __forceinline__ double plane_double(const EventLandau &ev, float x, int c) {
  if (c < FSM_NUM_OUTPUTS && !ev.fill_outputs) return 0.0;
  return plane_scaled(c) ? (double)x * (double)FSM_HBARC_F : (double)x;
}

// Float outputs of the function-level boundary (fsm_landau_device).
__device__ void cell_store_float(const EventLandau &ev, long is, const CellOut &r) {
  const long D = ev.dim;
  const unsigned i = (unsigned)is;
  if (ev.fe) fsm_gstore_at(ev.fe, i, r.e);
  if (ev.fu) {
#pragma unroll
    for (int c = 0; c < 4; ++c) fsm_gstore_at(ev.fu + c * D, i, r.u[c]);
  }
  if (ev.fp) fsm_gstore_at(ev.fp, i, r.p);
  if (ev.fPi) fsm_gstore_at(ev.fPi, i, r.Pi);
  if (ev.fpi) {
#pragma unroll
    for (int c = 0; c < 10; ++c) fsm_gstore_at(ev.fpi + c * D, i, r.pi[c]);
  }
}

// Landau frame and the outputs of one 2-D cell whose six stream-kernel
// components (T^{tau tau}, T^{tau x}, T^{tau y}, T^{xx}, T^{xy}, T^{yy}; the
// eta components are exact zeros) are in registers: the 2-D stream kernels'
// fused epilogue, the same arithmetic as the 3x3 Landau kernel on the same
// floats read back from memory.  is = the reference cell index; no baryon
// current on fused events.
__device__ __forceinline__ void landau2d_cell_store(const Event2D &s, const float t6[6], long is) {
  EventLandau ev = {};
  ev.eos_type = s.l_eos;
  ev.viscous = s.l_viscous;
  ev.fill_outputs = s.l_fill;
  ev.tau = s.tau;
  ev.inv_tau2 = s.l_inv_tau2;
  // The eta components as an opaque +0: known to the compiler as constants,
  // 0 - u^tau u^eta b could be folded to -(u^tau u^eta b), a -0 where the
  // Landau kernel, which reads them from memory, writes +0.
  float z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  const float t[10] = {t6[0], t6[1], t6[2], z, t6[3], t6[4], z, t6[5], z, z};
  CellOut r;
  landau_cell_from_t<3>(ev, t, 0, r);
#pragma unroll
  for (int c = 0; c < FSM_NUM_OUTPUTS; ++c)
    fsm_gstore_at(s.l_out + c * s.l_out_plane, (unsigned)is, plane_double(ev, cell_value(r, c), c));
}

#endif  // FSM_LANDAU_H_
