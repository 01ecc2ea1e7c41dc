// fsm_oracle.cpp -- CPU oracle for the free-streaming + Landau-matching step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke()
// entry of __graft_entry__.py and the cpu_baseline leg of bench.py load it;
// the product library (libfsm.so) never links, loads or calls it.
//
// It restates the reference algorithm (derekeverett/freestream-milne,
// src/FreeStream.cpp, src/LandauMatch.cpp, src/EquationOfState.cpp,
// src/FreestreamMilne.cpp) as plain C++ with OpenMP, keeping the reference's
// loop order, its unfused shiftedDensity[DIM][DIM_RAP][DIM_PHIP] array and its
// ten-pass quadrature.  One template, two instantiations:
//   R = float   the reference float CPU path (float storage and the same
//               float/double mixing of every expression, C++ <cmath>
//               overloads, so float arguments call the float functions);
//   R = double  the FP64 oracle (same algorithm, every quantity in double).
//
// Third-party arithmetic: the reference calls the GNU Scientific Library
// (version unpinned: CMakeLists.txt:9 find_package(GSL REQUIRED), Makefile:13
// -lgsl).  GSL is not installed here, so its two published algorithms are
// restated from the mathematics:
//   * gsl_interp2d_bicubic / gsl_spline2d (src/FreeStream.cpp:121-153, :240):
//     the bicubic Hermite patch whose knot derivatives z_x, z_y, z_xy are the
//     derivatives of 1-D natural cubic splines (gsl_interp_cspline) along
//     rows, columns, and along columns of z_x.  This is the tensor-product
//     natural cubic spline interpolant, which is unique.
//   * gsl_eigen_nonsymmv (src/LandauMatch.cpp:215-218): eigenvalues of the
//     real 4x4 matrix by Householder reduction to Hessenberg form and Francis
//     double-shift QR; eigenvectors of the real eigenvalues normalised to unit
//     Euclidean length.  The eigenvector here is the null vector of
//     (A - lambda I) by Gaussian elimination with complete pivoting (GSL
//     back-substitutes in the Schur form; both give the same unit vector up
//     to sign and rounding, and the reference's sign flip removes the sign).
//     GSL's balancing step is omitted (it changes rounding only).
// Parity pins for each part are listed in the tests that import this file
// (tests/test_oracle_*.py): scipy natural splines, numpy.linalg.eig, the
// Bessel closed forms of 2-D streaming of a Gaussian, the DTAU = 0 and
// uniform-density known answers, the pi.u = 0 / pi^mu_mu = 0 invariants, and
// (tests/test_oracle_pins.py) closed forms for 3-D streaming and the 3-D
// charge pass of affine profiles (trilinear interpolation is exact there),
// the 3-D charge pass at DTAU = 0, E_DEP_FS on a two-valued profile and the
// EOS_TYPE 2 fit against a double polyval of its coefficients.  Each of those
// pins is shown to fail on a planted mistake: oracle_set_mutation() switches
// one term of this file to a plausible wrong form (a trilinear corner, the
// sign of the eta shift, the Jacobian, a rapidity factor, a norm, ...); the
// product never calls it.

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/fsm.h"

namespace {

// ---------------------------------------------------------------------------
// 1-D natural cubic spline derivative at the knots (gsl_interp_cspline with
// natural end conditions).  Second-derivative coefficients c_i (= y''/2) solve
//   h_{i-1} c_{i-1} + 2 (h_{i-1}+h_i) c_i + h_i c_{i+1}
//       = 3 [ (y_{i+1}-y_i)/h_i - (y_i-y_{i-1})/h_{i-1} ],  c_0 = c_{n-1} = 0,
// and on interval i the slope is b_i + 2 c_i t + 3 d_i t^2 with
//   b_i = (y_{i+1}-y_i)/h_i - h_i (c_{i+1} + 2 c_i)/3,  d_i = (c_{i+1}-c_i)/(3 h_i).
// The derivative at knot i < n-1 is b_i; at the last knot it is the slope of
// interval n-2 at t = h_{n-2} (GSL's index lookup returns n-2 for x = x_max).
void natural_spline_knot_derivs(int n, const double *x, const double *y,
                                long ystride, double *dy, long dstride) {
  if (n < 2) {
    if (n == 1) dy[0] = 0.0;
    return;
  }
  std::vector<double> h(n - 1), c(n, 0.0);
  for (int i = 0; i < n - 1; ++i) h[i] = x[i + 1] - x[i];
  if (n > 2) {
    // Thomas algorithm on the interior unknowns c_1..c_{n-2}.
    int m = n - 2;
    std::vector<double> diag(m), off(m), rhs(m);
    for (int k = 0; k < m; ++k) {
      int i = k + 1;
      diag[k] = 2.0 * (h[i - 1] + h[i]);
      off[k] = h[i];  // coupling of c_i with c_{i+1} (symmetric)
      rhs[k] = 3.0 * ((y[(long)(i + 1) * ystride] - y[(long)i * ystride]) / h[i] -
                      (y[(long)i * ystride] - y[(long)(i - 1) * ystride]) / h[i - 1]);
    }
    for (int k = 1; k < m; ++k) {
      double w = off[k - 1] / diag[k - 1];
      diag[k] -= w * off[k - 1];
      rhs[k] -= w * rhs[k - 1];
    }
    c[m] = rhs[m - 1] / diag[m - 1];
    for (int k = m - 2; k >= 0; --k) c[k + 1] = (rhs[k] - off[k] * c[k + 2]) / diag[k];
  }
  for (int i = 0; i < n - 1; ++i) {
    double yi = y[(long)i * ystride], yn = y[(long)(i + 1) * ystride];
    double b = (yn - yi) / h[i] - h[i] * (c[i + 1] + 2.0 * c[i]) / 3.0;
    dy[(long)i * dstride] = b;
  }
  {
    int i = n - 2;
    double yi = y[(long)i * ystride], yn = y[(long)(i + 1) * ystride];
    double b = (yn - yi) / h[i] - h[i] * (c[i + 1] + 2.0 * c[i]) / 3.0;
    double d = (c[i + 1] - c[i]) / (3.0 * h[i]);
    dy[(long)(n - 1) * dstride] = b + 2.0 * c[i] * h[i] + 3.0 * d * h[i] * h[i];
  }
}

// Bicubic (tensor-product natural spline) on an nx x ny grid.  z is stored
// x-major: z[ix*ny + iy].
struct Spline2D {
  int nx = 0, ny = 0;
  std::vector<double> x, y, z, zx, zy, zxy;

  void init(int nx_, int ny_, const double *xv, const double *yv, const double *zv) {
    nx = nx_;
    ny = ny_;
    x.assign(xv, xv + nx);
    y.assign(yv, yv + ny);
    z.assign(zv, zv + (long)nx * ny);
    zx.assign((long)nx * ny, 0.0);
    zy.assign((long)nx * ny, 0.0);
    zxy.assign((long)nx * ny, 0.0);
    for (int iy = 0; iy < ny; ++iy)  // rows along x
      natural_spline_knot_derivs(nx, x.data(), z.data() + iy, ny, zx.data() + iy, ny);
    for (int ix = 0; ix < nx; ++ix) {  // columns along y
      natural_spline_knot_derivs(ny, y.data(), z.data() + (long)ix * ny, 1,
                                 zy.data() + (long)ix * ny, 1);
      natural_spline_knot_derivs(ny, y.data(), zx.data() + (long)ix * ny, 1,
                                 zxy.data() + (long)ix * ny, 1);
    }
  }

  // Largest i with a[i] <= v, clamped to [0, n-2] (interval lookup).
  static int locate(const std::vector<double> &a, double v) {
    int lo = 0, hi = (int)a.size() - 1;
    while (hi > lo + 1) {
      int mid = (lo + hi) / 2;
      if (a[mid] > v)
        hi = mid;
      else
        lo = mid;
    }
    return lo;
  }

  double eval(double xv, double yv) const {
    int i = locate(x, xv), j = locate(y, yv);
    double dx = x[i + 1] - x[i], dyv = y[j + 1] - y[j];
    double t = (xv - x[i]) / dx, u = (yv - y[j]) / dyv;
    // Cubic Hermite basis: value at 0, slope at 0, value at 1, slope at 1.
    double ht[4] = {(1 + 2 * t) * (1 - t) * (1 - t), t * (1 - t) * (1 - t),
                    t * t * (3 - 2 * t), t * t * (t - 1)};
    double hu[4] = {(1 + 2 * u) * (1 - u) * (1 - u), u * (1 - u) * (1 - u),
                    u * u * (3 - 2 * u), u * u * (u - 1)};
    double s = 0.0;
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) {
        long k = (long)(i + a) * ny + (j + b);
        double wt0 = ht[2 * a], wt1 = ht[2 * a + 1] * dx;
        double wu0 = hu[2 * b], wu1 = hu[2 * b + 1] * dyv;
        s += wt0 * wu0 * z[k] + wt1 * wu0 * zx[k] + wt0 * wu1 * zy[k] + wt1 * wu1 * zxy[k];
      }
    return s;
  }
};

// ---------------------------------------------------------------------------
// Real 4x4 eigen-decomposition (restatement of the published algorithm of
// gsl_eigen_nonsymmv: Hessenberg reduction + Francis double-shift QR).
struct Eig4 {
  double wr[4], wi[4];
  double v[4][4];  // v[row][k]: eigenvector k (valid where wi[k] == 0)
  int ok;
};

void hessenberg4(double H[4][4]) {
  for (int k = 0; k < 2; ++k) {
    double xn = 0.0;
    for (int i = k + 1; i < 4; ++i) xn += H[i][k] * H[i][k];
    xn = sqrt(xn);
    if (xn == 0.0) continue;
    double alpha = H[k + 1][k] > 0 ? -xn : xn;
    double vv[4] = {0, 0, 0, 0};
    for (int i = k + 1; i < 4; ++i) vv[i] = H[i][k];
    vv[k + 1] -= alpha;
    double vn = 0.0;
    for (int i = k + 1; i < 4; ++i) vn += vv[i] * vv[i];
    if (vn == 0.0) continue;
    double beta = 2.0 / vn;
    // H = P H P with P = I - beta v v^T
    for (int j = 0; j < 4; ++j) {
      double s = 0.0;
      for (int i = k + 1; i < 4; ++i) s += vv[i] * H[i][j];
      s *= beta;
      for (int i = k + 1; i < 4; ++i) H[i][j] -= s * vv[i];
    }
    for (int i = 0; i < 4; ++i) {
      double s = 0.0;
      for (int j = k + 1; j < 4; ++j) s += H[i][j] * vv[j];
      s *= beta;
      for (int j = k + 1; j < 4; ++j) H[i][j] -= s * vv[j];
    }
    for (int i = k + 2; i < 4; ++i) H[i][k] = 0.0;
  }
}

// Householder vector for (x, y, z) or (x, y) (nz = 2 or 3): P (x,y,z)^T = (*,0,0).
static void house(int nz, const double in[3], double vv[3], double *beta) {
  double xn = 0.0;
  for (int i = 0; i < nz; ++i) xn += in[i] * in[i];
  xn = sqrt(xn);
  for (int i = 0; i < 3; ++i) vv[i] = (i < nz) ? in[i] : 0.0;
  if (xn == 0.0) {
    *beta = 0.0;
    return;
  }
  double alpha = in[0] > 0 ? -xn : xn;
  vv[0] -= alpha;
  double vn = 0.0;
  for (int i = 0; i < nz; ++i) vn += vv[i] * vv[i];
  *beta = vn == 0.0 ? 0.0 : 2.0 / vn;
}

static void apply_house(double H[4][4], int k, int nz, const double vv[3], double beta,
                        int l, int m) {
  if (beta == 0.0) return;
  for (int j = l; j <= m; ++j) {
    double s = 0.0;
    for (int i = 0; i < nz; ++i) s += vv[i] * H[k + i][j];
    s *= beta;
    for (int i = 0; i < nz; ++i) H[k + i][j] -= s * vv[i];
  }
  for (int i = l; i <= m; ++i) {
    double s = 0.0;
    for (int j = 0; j < nz; ++j) s += H[i][k + j] * vv[j];
    s *= beta;
    for (int j = 0; j < nz; ++j) H[i][k + j] -= s * vv[j];
  }
}

// Francis double-shift QR step on the active window [l, m] (m - l >= 2).
static void francis_step(double H[4][4], int l, int m, int exceptional) {
  double s, t;
  if (exceptional) {
    double w = fabs(H[m][m - 1]) + fabs(H[m - 1][m - 2]);
    s = 1.5 * w + H[m][m];
    t = w * w;
    s = 2.0 * s;  // shifts at roughly H[m][m] +/- w
    t = (H[m][m] + 0.75 * w) * (H[m][m] + 0.75 * w);
  } else {
    s = H[m - 1][m - 1] + H[m][m];
    t = H[m - 1][m - 1] * H[m][m] - H[m - 1][m] * H[m][m - 1];
  }
  double xyz[3];
  xyz[0] = H[l][l] * H[l][l] + H[l][l + 1] * H[l + 1][l] - s * H[l][l] + t;
  xyz[1] = H[l + 1][l] * (H[l][l] + H[l + 1][l + 1] - s);
  xyz[2] = H[l + 1][l] * H[l + 2][l + 1];
  for (int k = l; k <= m - 2; ++k) {
    double vv[3], beta;
    house(3, xyz, vv, &beta);
    apply_house(H, k, 3, vv, beta, l, m);
    xyz[0] = H[k + 1][k];
    xyz[1] = H[k + 2][k];
    if (k < m - 2) xyz[2] = H[k + 3][k];
  }
  {
    double vv[3], beta;
    house(2, xyz, vv, &beta);
    apply_house(H, m - 1, 2, vv, beta, l, m);
  }
}

static void eig_values4(double A[4][4], double wr[4], double wi[4], int *ok) {
  double H[4][4];
  memcpy(H, A, sizeof(H));
  hessenberg4(H);
  const double eps = 2.220446049250313e-16;
  int m = 3, iter = 0, total = 0;
  *ok = 1;
  while (m >= 0) {
    int l = m;
    while (l > 0) {
      double sc = fabs(H[l - 1][l - 1]) + fabs(H[l][l]);
      if (sc == 0.0) sc = 1.0;
      if (fabs(H[l][l - 1]) <= eps * sc) {
        H[l][l - 1] = 0.0;
        break;
      }
      --l;
    }
    if (l == m) {
      wr[m] = H[m][m];
      wi[m] = 0.0;
      --m;
      iter = 0;
    } else if (l == m - 1) {
      double a = H[m - 1][m - 1], b = H[m - 1][m], c = H[m][m - 1], d = H[m][m];
      double p = 0.5 * (a - d), q = p * p + b * c;
      double mid = 0.5 * (a + d);
      if (q >= 0.0) {
        double r = sqrt(q);
        // larger-magnitude root first, the other from the product (stable)
        double big = mid + (mid >= 0 ? r : -r);
        double prod = a * d - b * c;
        double other = (big != 0.0) ? prod / big : mid - (mid >= 0 ? r : -r);
        wr[m - 1] = big;
        wr[m] = other;
        wi[m - 1] = wi[m] = 0.0;
      } else {
        double r = sqrt(-q);
        wr[m - 1] = wr[m] = mid;
        wi[m - 1] = r;
        wi[m] = -r;
      }
      m -= 2;
      iter = 0;
    } else {
      ++iter;
      ++total;
      if (total > 200) {
        *ok = 0;
        for (int i = 0; i <= m; ++i) {
          wr[i] = H[i][i];
          wi[i] = 0.0;
        }
        return;
      }
      francis_step(H, l, m, (iter % 11) == 10);
    }
  }
}

// Null vector of (A - lam I) by Gaussian elimination with complete pivoting.
static void null_vector4(double A[4][4], double lam, double out[4]) {
  double M[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) M[i][j] = A[i][j] - (i == j ? lam : 0.0);
  int colp[4] = {0, 1, 2, 3};
  for (int k = 0; k < 3; ++k) {
    int pi = k, pj = k;
    double best = -1.0;
    for (int i = k; i < 4; ++i)
      for (int j = k; j < 4; ++j)
        if (fabs(M[i][j]) > best) {
          best = fabs(M[i][j]);
          pi = i;
          pj = j;
        }
    if (pi != k)
      for (int j = 0; j < 4; ++j) std::swap(M[k][j], M[pi][j]);
    if (pj != k) {
      for (int i = 0; i < 4; ++i) std::swap(M[i][k], M[i][pj]);
      std::swap(colp[k], colp[pj]);
    }
    if (M[k][k] == 0.0) continue;
    for (int i = k + 1; i < 4; ++i) {
      double f = M[i][k] / M[k][k];
      for (int j = k; j < 4; ++j) M[i][j] -= f * M[k][j];
    }
  }
  double yv[4];
  yv[3] = 1.0;
  for (int k = 2; k >= 0; --k) {
    double s = 0.0;
    for (int j = k + 1; j < 4; ++j) s += M[k][j] * yv[j];
    yv[k] = (M[k][k] != 0.0) ? -s / M[k][k] : 0.0;
  }
  double nrm = 0.0;
  for (int k = 0; k < 4; ++k) nrm += yv[k] * yv[k];
  nrm = sqrt(nrm);
  for (int k = 0; k < 4; ++k) out[colp[k]] = yv[k] / nrm;
}

void eig4(double A[4][4], Eig4 *r) {
  eig_values4(A, r->wr, r->wi, &r->ok);
  for (int k = 0; k < 4; ++k) {
    double vec[4] = {0, 0, 0, 0};
    if (r->wi[k] == 0.0) null_vector4(A, r->wr[k], vec);
    for (int i = 0; i < 4; ++i) r->v[i][k] = vec[i];
  }
}

// ---------------------------------------------------------------------------
// Landau matching of one cell: src/LandauMatch.cpp:163-291.  T in, e and u
// out (float storage as in the reference, or double for the FP64 oracle).
template <typename R>
void landau_cell(const R T10[10], float TAU, R *e_out, R u_out[4]) {
  double A[4][4];
  const double T[10] = {(double)T10[0], (double)T10[1], (double)T10[2], (double)T10[3],
                        (double)T10[4], (double)T10[5], (double)T10[6], (double)T10[7],
                        (double)T10[8], (double)T10[9]};
  // symmetric fill :177-192
  A[0][0] = T[0]; A[0][1] = T[1]; A[0][2] = T[2]; A[0][3] = T[3];
  A[1][1] = T[4]; A[1][2] = T[5]; A[1][3] = T[6];
  A[2][2] = T[7]; A[2][3] = T[8]; A[3][3] = T[9];
  A[1][0] = T[1]; A[2][0] = T[2]; A[3][0] = T[3];
  A[2][1] = T[5]; A[3][1] = T[6]; A[3][2] = T[8];
  // element-wise metric factors :197-214, g_(nu) = (1, -1, -1, -TAU^2)
  const double gcol[4] = {1.0, -1.0, -1.0, -1.0 * TAU * TAU};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) A[i][j] *= gcol[j];
  Eig4 ev;
  eig4(A, &ev);
  const float gamma_max = 100.0;
  // TAU*TAU is a float product in the reference (:235,237); the FP64 oracle
  // forms it in double.
  const double tau2 = (double)((R)TAU * (R)TAU);
  int found = 0;
  for (int i = 0; i < 4; ++i) {  // :222-281, last match wins
    double re = ev.wr[i], im = ev.wi[i];
    if (re > 0.0 && fabs(im) < fabs(re) * 1.0e-30) {
      double v0 = ev.v[0][i], v1 = ev.v[1][i], v2 = ev.v[2][i], v3 = ev.v[3][i];
      if ((2.0 * v0 * v0 - 1.0 - (v3 * v3 * (tau2 - 1.0))) > 0) {
        double ml = v0 * v0 - (v1 * v1 + v2 * v2 + tau2 * v3 * v3);
        double factor = 1.0 / sqrt(ml);
        if (v0 < 0) factor = -factor;
        if ((v0 * factor) < gamma_max) {
          found = 1;
          *e_out = (R)re;
          u_out[0] = (R)(v0 * factor);
          u_out[1] = (R)(v1 * factor);
          u_out[2] = (R)(v2 * factor);
          u_out[3] = (R)(v3 * factor);
        }
      }
    }
  }
  if (!found) {  // :283-291
    *e_out = (R)1.0e-3;
    u_out[0] = (R)1.0;
    u_out[1] = u_out[2] = u_out[3] = (R)0.0;
  }
}

// ---------------------------------------------------------------------------
const float kHbarc = 0.197326938;  // src/FreestreamMilne.cpp:141 (float)

// Planted mistakes for the tests that show each pin catches an error in the
// function it pins (oracle_set_mutation; 0 = none, the only value outside
// those tests).  Each id changes one term of one function:
enum Mutation {
  MUT_NONE = 0,
  MUT_TRILINEAR_CORNER = 1,  // freeStream 3-D: corners (1,0,0) and (0,1,0) swapped
  MUT_ETA_NEW_SIGN = 2,      // freeStream 3-D: eta_new = asinh(...) - rap
  MUT_SHIFT_NO_TAU0 = 3,     // freeStream 3-D: D without the -TAU0 cosh(rap - eta) term
  MUT_JACOBIAN_ONE_COS = 4,  // stress tensor 3-D: Jacobian with one cos factor
  MUT_3D_OVER_TAU = 5,       // stress tensor 3-D: acc * d_phip / TAU as in 2-D
  MUT_CHARGE_WIDTH = 6,      // convertInitialChargeDensity: exp(-d^2 / (2 sigma_B^2))
  MUT_CHARGE_NORM = 7,       // convertInitialChargeDensity: n without exp(sigma_B^2 / 2)
  MUT_EDEP_NO_HBARC = 8,     // getEnergyDependentTau: e_T not multiplied by hbarc
  MUT_EOS2_SWAP = 9,         // calculatePressure EOS 2: a5 and a6 swapped
  MUT_HYPERTRIG_TAU = 10,    // hypertrig row 3: th instead of th / TAU
  MUT_ETA_MASK_LOW = 11,     // freeStream 3-D: keeps eta floor index 0
  MUT_RAP_FACTOR_COSH = 12,  // convertInitialDensity 3-D: cosh instead of cosh^2
  MUT_BARYON_NO_DRAP = 13,   // calculateBaryonCurrent: J without the DRAP factor
};
int g_mutation = MUT_NONE;
inline bool mut(int m) { return g_mutation == m; }

// Rapidity rule of the 3-D path.  0: the reference's tan grid (the only rule
// the reference has).  1: the Gauss-Hermite rule of the library's
// FSM_RAPIDITY_GAUSS_HERMITE extension, nodes eta + sqrt(2) SIGMA x_k and
// weights sqrt(2) SIGMA W_k with the density keeping norm * cosh^2(eta - y);
// x_k, W_k are handed in by oracle_set_rapidity_rule (the tests take them
// from numpy, independently of the library's own rule).
int g_rap_rule = 0;
std::vector<double> g_gh_x, g_gh_w;

// Rapidity node irap of a cell at eta (src/FreeStream.cpp:174-175 and
// :337-340, src/LandauMatch.cpp:37-38 under the tan grid).
template <typename R>
R rap_node(R eta, int irap, int DIM_RAP, double sigma) {
  if (g_rap_rule == 1) return (R)((double)eta + sqrt(2.0) * sigma * g_gh_x[irap]);
  R w = (R)(-.9975 + (double)(R)irap * (1.995 / (double)(R)(DIM_RAP - 1)));
  return (R)((double)eta + tan((M_PI / 2.0) * (double)w));
}

// EquationOfState.cpp:5-80, one cell.
template <typename R>
R pressure_of(R e, int eos) {
  if (eos == 1) return (R)(e / 3.0);
  if (eos == 2) {
    R a[13] = {(R)-0.25181736420168666, (R)9737.845799644809, (R)1.077580993288114e6,
                     (R)3.1729694865420084e6, (R)1.6357487344679043e6, (R)334334.4309240126,
                     (R)41913.439282708554, (R)6340.448389300905, (R)141.5073484468774,
                     (R)0.7158279081255019, (R)0.0009417586777847889,
                     (R)3.1188455176941583e-7, (R)1.9531729608963267e-11};
    const R b[13] = {(R)45829.44617893836, (R)4.0574329080826794e6, (R)2.0931169138134286e7,
                     (R)1.3512402226067686e7, (R)1.7851642641834426e6, (R)278581.2989342773,
                     (R)26452.34905933697, (R)499.04919730607065, (R)2.3405487982094204,
                     (R)0.002962497695527404, (R)9.601103399348206e-7,
                     (R)5.928138360995685e-11, (R)3.2581066229887368e-18};
    if (mut(MUT_EOS2_SWAP)) std::swap(a[5], a[6]);
    R ep[13];
    ep[0] = 1;
    ep[1] = e;
    for (int k = 2; k < 13; ++k) ep[k] = ep[k - 1] * e;
    R na = a[0], nb = b[0];
    for (int k = 1; k < 13; ++k) {  // innermost fmaf first, as :555-556
      na = std::fma(a[k], ep[k], na);
      nb = std::fma(b[k], ep[k], nb);
    }
    // the reference nests fmaf(a12,e12, fmaf(a11,e11, ... fmaf(a1,e1,a0)))
    return na / nb;
  }
  return (R)0.0;
}

template <typename R>
struct Pipeline {
  const fsm_params &P;
  explicit Pipeline(const fsm_params &p) : P(p) {}

  // src/FreestreamMilne.cpp:267
  void ingest(const float *v, std::vector<R> &e0) const {
    const R lower_tolerance = (R)1.0e-7;
    const R rescale = (R)1.0;
    e0.resize(P.DIM);
    // the vector input (IC_ENERGY = 5) in GeV/fm^3; the other initial
    // conditions hand their fm^-4 profile over as is (:231-256)
    for (int i = 0; i < P.DIM; ++i)
      e0[i] = P.IC_ENERGY == 5 ? (R)v[i] * rescale / (R)kHbarc + lower_tolerance : (R)v[i];
  }

  // src/FreeStream.cpp:289-350 (density[DIM][DIM_RAP], flat)
  void convert_density(const std::vector<R> &e0, std::vector<R> &dens) const {
    const int DIM = P.DIM, DIM_ETA = P.DIM_ETA, DIM_Y = P.DIM_Y, DIM_RAP = P.DIM_RAP;
    const R SIGMA = (R)P.SIGMA, DETA = (R)P.DETA;
    dens.assign((size_t)DIM * DIM_RAP, (R)0);
    R n = (R)(sqrt(M_PI / 2.0) * (double)SIGMA * (1.0 + exp(2.0 * (double)SIGMA * (double)SIGMA)));
    R norm_factor = (R)(1.0 / (2.0 * M_PI * (double)n));
    R etamin = (R)((-1.0) * ((double)(R)(DIM_ETA - 1) / 2.0) * (double)DETA);
    if (DIM_ETA == 1) {
      for (int is = 0; is < DIM; ++is)
        for (int irap = 0; irap < DIM_RAP; ++irap)
          dens[(size_t)is * DIM_RAP + irap] = (R)((double)e0[is] / (2.0 * M_PI));
    } else {
      for (int is = 0; is < DIM; ++is) {
        int ix = is / (DIM_Y * DIM_ETA);
        int iy = (is - (DIM_Y * DIM_ETA * ix)) / DIM_ETA;
        int ieta = is - (DIM_Y * DIM_ETA * ix) - (DIM_ETA * iy);
        R eta = (R)ieta * DETA + etamin;
        for (int irap = 0; irap < DIM_RAP; ++irap) {
          R rap = rap_node<R>(eta, irap, DIM_RAP, (double)SIGMA);
          R d = eta - rap;
          const double gauss =
              g_rap_rule == 1 ? 1.0 : exp((-1.0) * (double)d * (double)d / (2.0 * (double)SIGMA * (double)SIGMA));
          R rap_factor = (R)((double)(std::cosh(d) * (mut(MUT_RAP_FACTOR_COSH) ? (R)1 : std::cosh(d))) * gauss);
          dens[(size_t)is * DIM_RAP + irap] = e0[is] * norm_factor * rap_factor;
        }
      }
    }
  }

  // src/FreeStream.cpp:353-387
  void convert_charge(const std::vector<R> &n0, std::vector<R> &dens) const {
    const int DIM = P.DIM, DIM_ETA = P.DIM_ETA, DIM_Y = P.DIM_Y, DIM_RAP = P.DIM_RAP;
    const R SIGMA_B = (R)P.SIGMA_B, DRAP = (R)P.DRAP, DETA = (R)P.DETA;
    dens.assign((size_t)DIM * DIM_RAP, (R)0);
    // SIGMA_B * SIGMA_B is a float product in the reference (float operands)
    const R sb2 = SIGMA_B * SIGMA_B;
    R n = (R)(sqrt(2.0 * M_PI) * (double)SIGMA_B * (mut(MUT_CHARGE_NORM) ? 1.0 : exp((double)sb2 / 2.0)));
    R norm_factor = (R)(1.0 / (2.0 * M_PI * (double)n));
    R rapmin = (R)((-1.0) * ((double)(R)(DIM_RAP - 1) / 2.0) * (double)DRAP);
    R etamin = (R)((-1.0) * ((double)(R)(DIM_ETA - 1) / 2.0) * (double)DETA);
    for (int is = 0; is < DIM; ++is) {
      int ix = is / (DIM_Y * DIM_ETA);
      int iy = (is - (DIM_Y * DIM_ETA * ix)) / DIM_ETA;
      int ieta = is - (DIM_Y * DIM_ETA * ix) - (DIM_ETA * iy);
      R eta = (R)ieta * DETA + etamin;
      for (int irap = 0; irap < DIM_RAP; ++irap) {
        R rap = (R)irap * DRAP + rapmin;
        R d = eta - rap;
        R rap_factor = (R)((double)std::cosh(d) *
                           exp((-1.0) * (double)d * (double)d / ((mut(MUT_CHARGE_WIDTH) ? 2.0 : 1.0) * (double)sb2)));
        dens[(size_t)is * DIM_RAP + irap] = n0[is] * norm_factor * rap_factor;
      }
    }
  }

  // src/FreeStream.cpp:97-284; shifted[DIM][DIM_RAP][DIM_PHIP] flat.
  void free_stream(const std::vector<R> &dens, std::vector<R> &shifted) const {
    const int DIM_X = P.DIM_X, DIM_Y = P.DIM_Y, DIM_ETA = P.DIM_ETA, DIM_RAP = P.DIM_RAP,
              DIM_PHIP = P.DIM_PHIP, DIM = P.DIM;
    const R DX = (R)P.DX, DY = (R)P.DY, DETA = (R)P.DETA, TAU0 = (R)P.TAU0, TAU = (R)P.TAU,
            DTAU = (R)P.DTAU;
    R xmin = (R)((-1.0) * ((double)(R)(DIM_X - 1) / 2.0) * (double)DX);
    R ymin = (R)((-1.0) * ((double)(R)(DIM_Y - 1) / 2.0) * (double)DY);
    R etamin = (R)((-1.0) * ((double)(R)(DIM_ETA - 1) / 2.0) * (double)DETA);
    shifted.assign((size_t)DIM * DIM_RAP * DIM_PHIP, (R)0);

    // Spline setup :121-153 (x_vals/y_vals are the grid coordinates as
    // computed by getX/getY, promoted to double; density_vals from irap 0).
    Spline2D spline;
    std::vector<double> xv(DIM_X), yv(DIM_Y), zv((size_t)DIM_X * DIM_Y);
    if (DIM_ETA == 1) {
      for (int ix = 0; ix < DIM_X; ++ix) xv[ix] = (double)((R)ix * DX + xmin);
      for (int iy = 0; iy < DIM_Y; ++iy) yv[iy] = (double)((R)iy * DY + ymin);
      for (int ix = 0; ix < DIM_X; ++ix)
        for (int iy = 0; iy < DIM_Y; ++iy)
          zv[(size_t)ix * DIM_Y + iy] = (double)dens[((size_t)ix * DIM_Y + iy) * DIM_RAP + 0];
      spline.init(DIM_X, DIM_Y, xv.data(), yv.data(), zv.data());
    }
    const double x_min_interp = DIM_ETA == 1 ? xv[0] : 0.0;
    const double y_min_interp = DIM_ETA == 1 ? yv[0] : 0.0;
    const double x_max_interp = DIM_ETA == 1 ? xv[DIM_X - 1] : 0.0;
    const double y_max_interp = DIM_ETA == 1 ? yv[DIM_Y - 1] : 0.0;

#pragma omp parallel for schedule(static)
    for (int is = 0; is < DIM; ++is) {
      int ix = is / (DIM_Y * DIM_ETA);
      int iy = (is - (DIM_Y * DIM_ETA * ix)) / DIM_ETA;
      int ieta = is - (DIM_Y * DIM_ETA * ix) - (DIM_ETA * iy);
      R x = (R)ix * DX + xmin;
      R y = (R)iy * DY + ymin;
      R eta = (R)ieta * DETA + etamin;
      for (int irap = 0; irap < DIM_RAP; ++irap) {
        R rap = rap_node<R>(eta, irap, DIM_RAP, P.SIGMA);
        for (int iphip = 0; iphip < DIM_PHIP; ++iphip) {
          R phip = (R)((double)(R)iphip * (2.0 * M_PI) / (double)(R)DIM_PHIP);
          R eta_new, x_new, y_new;
          if (DIM_ETA == 1) {
            eta_new = 0.0;
            x_new = x - std::cos(phip) * DTAU;
            y_new = y - std::sin(phip) * DTAU;
          } else {
            eta_new = std::asinh((TAU / TAU0) * std::sinh(eta - rap)) + (mut(MUT_ETA_NEW_SIGN) ? -rap : rap);
            const R t0 = mut(MUT_SHIFT_NO_TAU0) ? (R)0 : TAU0;
            x_new = x - std::cos(phip) * (TAU * std::cosh(rap - eta_new) - t0 * std::cosh(rap - eta));
            y_new = y - std::sin(phip) * (TAU * std::cosh(rap - eta_new) - t0 * std::cosh(rap - eta));
          }
          R ix_new = (x_new - xmin) / DX;
          R iy_new = (y_new - ymin) / DY;
          R ieta_new = (DIM_ETA > 1) ? (eta_new - etamin) / DETA : (R)0;
          R ix_new_f = std::floor(ix_new);
          R iy_new_f = std::floor(iy_new);
          R ieta_new_f = std::floor(ieta_new);
          R x_frac = ix_new - ix_new_f;
          R y_frac = iy_new - iy_new_f;
          R eta_frac = ieta_new - ieta_new_f;
          if ((ix_new_f >= 1) && (ix_new_f < DIM_X - 1) && (iy_new_f >= 1) && (iy_new_f < DIM_Y - 1)) {
            R interp = 0.0;
            if (DIM_ETA == 1) {
              int i0 = (int)ix_new_f, j0 = (int)iy_new_f;
              R a11 = dens[((size_t)(DIM_Y * i0 + j0)) * DIM_RAP + irap];
              R a21 = dens[((size_t)(DIM_Y * (i0 + 1) + j0)) * DIM_RAP + irap];
              R a12 = dens[((size_t)(DIM_Y * i0 + j0 + 1)) * DIM_RAP + irap];
              R a22 = dens[((size_t)(DIM_Y * (i0 + 1) + j0 + 1)) * DIM_RAP + irap];
              R interp_l = ((1 - x_frac) * (1 - y_frac) * a11) + ((x_frac) * (1 - y_frac) * a21) +
                           ((1 - x_frac) * (y_frac) * a12) + ((x_frac) * (y_frac) * a22);
              R interp_c = 0.0;
              if ((x_min_interp < (double)x_new) && ((double)x_new < x_max_interp) &&
                  (y_min_interp < (double)y_new) && ((double)y_new < y_max_interp))
                interp_c = (R)spline.eval((double)x_new, (double)y_new);
              interp = (interp_c > 0.0) ? interp_c : interp_l;
            } else if ((ieta_new_f >= (mut(MUT_ETA_MASK_LOW) ? 0 : 1)) && (ieta_new_f < DIM_ETA - 1)) {
              int i0 = (int)ix_new_f, j0 = (int)iy_new_f, k0 = (int)ieta_new_f;
              auto D = [&](int a, int b, int c) {
                return dens[((size_t)(DIM_Y * DIM_ETA * (i0 + a) + DIM_ETA * (j0 + b) + (k0 + c))) * DIM_RAP + irap];
              };
              R x0 = x_frac, x1 = y_frac, x2 = eta_frac;
              const R c100 = mut(MUT_TRILINEAR_CORNER) ? D(0, 1, 0) : D(1, 0, 0);
              const R c010 = mut(MUT_TRILINEAR_CORNER) ? D(1, 0, 0) : D(0, 1, 0);
              interp = ((1 - x0) * (1 - x1) * (1 - x2) * D(0, 0, 0)) + ((x0) * (1 - x1) * (1 - x2) * c100) +
                       ((1 - x0) * (x1) * (1 - x2) * c010) + ((1 - x0) * (1 - x1) * (x2) * D(0, 0, 1)) +
                       ((x0) * (x1) * (1 - x2) * D(1, 1, 0)) + ((x0) * (1 - x1) * (x2) * D(1, 0, 1)) +
                       ((1 - x0) * (x1) * (x2) * D(0, 1, 1)) + ((x0) * (x1) * (x2) * D(1, 1, 1));
            }
            shifted[((size_t)is * DIM_RAP + irap) * DIM_PHIP + iphip] = interp;
          }
        }
      }
    }
  }

  // src/LandauMatch.cpp:14-59; H[10][DIM_RAP][DIM_PHIP][DIM_ETA] flat.
  void hypertrig(std::vector<R> &H) const {
    const int DIM_RAP = P.DIM_RAP, DIM_PHIP = P.DIM_PHIP, DIM_ETA = P.DIM_ETA;
    const R DETA = (R)P.DETA, TAU = (R)P.TAU;
    R etamin = (R)((-1.0) * ((double)(R)(DIM_ETA - 1) / 2.0) * (double)DETA);
    H.assign((size_t)10 * DIM_RAP * DIM_PHIP * DIM_ETA, (R)0);
    auto at = [&](int v, int k, int j, int ie) -> R & {
      return H[(((size_t)v * DIM_RAP + k) * DIM_PHIP + j) * DIM_ETA + ie];
    };
    for (int irap = 0; irap < DIM_RAP; ++irap)
      for (int iphip = 0; iphip < DIM_PHIP; ++iphip) {
        R phip = (R)((double)(R)iphip * (2.0 * M_PI) / (double)(R)DIM_PHIP);
        for (int ieta = 0; ieta < DIM_ETA; ++ieta) {
          R eta = (R)ieta * DETA + etamin;
          if (DIM_ETA == 1) eta = 0.0;
          R rap = rap_node<R>(eta, irap, DIM_RAP, P.SIGMA);
          if (DIM_ETA == 1) rap = 0.0;
          R c = std::cos(phip), s = std::sin(phip);
          R ch = std::cosh(rap - eta), th = std::tanh(rap - eta);
          at(0, irap, iphip, ieta) = 1.0;
          at(1, irap, iphip, ieta) = c / ch;
          at(2, irap, iphip, ieta) = s / ch;
          at(3, irap, iphip, ieta) = (R)((mut(MUT_HYPERTRIG_TAU) ? 1.0 : 1.0 / (double)TAU) * (double)th);
          at(4, irap, iphip, ieta) = (c * c) / (ch * ch);
          at(5, irap, iphip, ieta) = (c * s) / (ch * ch);
          at(6, irap, iphip, ieta) = (R)((1.0 / (double)TAU) * (double)(c * th) / (double)ch);
          at(7, irap, iphip, ieta) = (s * s) / (ch * ch);
          at(8, irap, iphip, ieta) = (R)((1.0 / (double)TAU) * (double)(s * th) / (double)ch);
          at(9, irap, iphip, ieta) = (R)((1.0 / (double)(TAU * TAU)) * (double)th * (double)th);
        }
      }
  }

  // src/LandauMatch.cpp:61-108; T[10][DIM] flat (ten serial passes).
  void stress_tensor(const std::vector<R> &shifted, const std::vector<R> &H, std::vector<R> &T) const {
    const int DIM_Y = P.DIM_Y, DIM_ETA = P.DIM_ETA, DIM_RAP = P.DIM_RAP, DIM_PHIP = P.DIM_PHIP,
              DIM = P.DIM;
    const R TAU = (R)P.TAU;
    R d_phip = (R)((2.0 * M_PI) / (double)(R)DIM_PHIP);
    T.assign((size_t)10 * DIM, (R)0);
    for (int ivar = 0; ivar < 10; ++ivar) {
#pragma omp parallel for schedule(static)
      for (int is = 0; is < DIM; ++is) {
        int ix = is / (DIM_Y * DIM_ETA);
        int iy = (is - (DIM_Y * DIM_ETA * ix)) / DIM_ETA;
        int ieta = is - (DIM_Y * DIM_ETA * ix) - (DIM_ETA * iy);
        R acc = T[(size_t)ivar * DIM + is];
        for (int irap = 0; irap < DIM_RAP; ++irap) {
          R w = (R)(-.9975 + (double)(R)irap * (1.995 / (double)(R)(DIM_RAP - 1)));
          R jacobian = g_rap_rule == 1
                           ? (R)(sqrt(2.0) * P.SIGMA * g_gh_w[irap])
                           : (R)((M_PI / 2.0) / cos((M_PI / 2.0) * (double)w) /
                                 (mut(MUT_JACOBIAN_ONE_COS) ? 1.0 : cos((M_PI / 2.0) * (double)w)) *
                                 (1.995 / (double)(R)(DIM_RAP - 1)));
          for (int iphip = 0; iphip < DIM_PHIP; ++iphip) {
            R s = shifted[((size_t)is * DIM_RAP + irap) * DIM_PHIP + iphip];
            R h = H[(((size_t)ivar * DIM_RAP + irap) * DIM_PHIP + iphip) * DIM_ETA + ieta];
            if (DIM_ETA == 1)
              acc += s * h;
            else
              acc += s * h * jacobian;
          }
        }
        if (DIM_ETA == 1 || mut(MUT_3D_OVER_TAU))
          acc = acc * d_phip / TAU;
        else
          acc = acc * d_phip;
        T[(size_t)ivar * DIM + is] = acc;
      }
    }
  }

  // src/LandauMatch.cpp:110-143; J[4][DIM].
  void baryon_current(const std::vector<R> &shifted, const std::vector<R> &H, std::vector<R> &J) const {
    const int DIM_Y = P.DIM_Y, DIM_ETA = P.DIM_ETA, DIM_RAP = P.DIM_RAP, DIM_PHIP = P.DIM_PHIP,
              DIM = P.DIM;
    const R DRAP = (R)P.DRAP;
    R d_phip = (R)((2.0 * M_PI) / (double)(R)DIM_PHIP);
    J.assign((size_t)4 * DIM, (R)0);
    for (int ivar = 0; ivar < 4; ++ivar) {
#pragma omp parallel for schedule(static)
      for (int is = 0; is < DIM; ++is) {
        int ix = is / (DIM_Y * DIM_ETA);
        int iy = (is - (DIM_Y * DIM_ETA * ix)) / DIM_ETA;
        int ieta = is - (DIM_Y * DIM_ETA * ix) - (DIM_ETA * iy);
        R acc = 0;
        for (int irap = 0; irap < DIM_RAP; ++irap)
          for (int iphip = 0; iphip < DIM_PHIP; ++iphip)
            acc += shifted[((size_t)is * DIM_RAP + irap) * DIM_PHIP + iphip] *
                   H[(((size_t)ivar * DIM_RAP + irap) * DIM_PHIP + iphip) * DIM_ETA + ieta];
        J[(size_t)ivar * DIM + is] = acc * (mut(MUT_BARYON_NO_DRAP) ? (R)1 : DRAP) * d_phip;
      }
    }
  }
};

// calculateBulkPressure (src/LandauMatch.cpp:332-344), then
// calculateShearViscTensor (:345-367), on planes of DIM values.
template <typename R>
void viscous_stages(int DIM, R TAU, const R *T, const R *e, const R *u, const R *p, R *Pi, R *pi) {
  for (int is = 0; is < DIM; ++is) {
    R a = T[0 * (size_t)DIM + is] - T[4 * (size_t)DIM + is] - T[7 * (size_t)DIM + is] -
          TAU * TAU * T[9 * (size_t)DIM + is];
    Pi[is] = (R)((-1.0 / 3.0) * (double)(a - e[is]) - (double)p[is]);
  }
  for (int is = 0; is < DIM; ++is) {
    R b = e[is] + p[is] + Pi[is];
    R c = p[is] + Pi[is];
    const R *uc[4] = {&u[0 * (size_t)DIM + is], &u[1 * (size_t)DIM + is], &u[2 * (size_t)DIM + is],
                      &u[3 * (size_t)DIM + is]};
    auto Tc = [&](int k) { return T[(size_t)k * DIM + is]; };
    pi[0 * (size_t)DIM + is] = Tc(0) - *uc[0] * *uc[0] * b + c;
    pi[1 * (size_t)DIM + is] = Tc(1) - *uc[0] * *uc[1] * b;
    pi[2 * (size_t)DIM + is] = Tc(2) - *uc[0] * *uc[2] * b;
    pi[3 * (size_t)DIM + is] = Tc(3) - *uc[0] * *uc[3] * b;
    pi[4 * (size_t)DIM + is] = Tc(4) - *uc[1] * *uc[1] * b - c;
    pi[5 * (size_t)DIM + is] = Tc(5) - *uc[1] * *uc[2] * b;
    pi[6 * (size_t)DIM + is] = Tc(6) - *uc[1] * *uc[3] * b;
    pi[7 * (size_t)DIM + is] = Tc(7) - *uc[2] * *uc[2] * b - c;
    pi[8 * (size_t)DIM + is] = Tc(8) - *uc[2] * *uc[3] * b;
    pi[9 * (size_t)DIM + is] = (R)((double)(Tc(9) - *uc[3] * *uc[3] * b) - (double)c * (1.0 / (double)(TAU * TAU)));
  }
}

template <typename R>
int run_pipeline(const fsm_params &P, const float *e_in, const float *nB_in, double *out17,
                 double *T_out, double *extra, double *stage_seconds) {
  Pipeline<R> pl(P);
  const int DIM = P.DIM;
  const R TAU = (R)P.TAU;
  auto now = []() {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
  };
  double t0 = now();
  std::vector<R> e0, dens, shifted, H, T;
  pl.ingest(e_in, e0);
  pl.convert_density(e0, dens);
  std::vector<R>().swap(e0);
  pl.free_stream(dens, shifted);
  double t1 = now();
  std::vector<R>().swap(dens);
  pl.hypertrig(H);
  pl.stress_tensor(shifted, H, T);
  std::vector<R>().swap(shifted);
  double t2 = now();

  std::vector<R> J;
  if (P.BARYON && nB_in) {
    std::vector<R> n0(nB_in, nB_in + DIM), cd, sc;
    pl.convert_charge(n0, cd);
    pl.free_stream(cd, sc);
    pl.baryon_current(sc, H, J);
  }

  std::vector<R> e(DIM), u((size_t)4 * DIM), p(DIM, (R)0), Pi(DIM, (R)0), pi((size_t)10 * DIM, (R)0);
#pragma omp parallel for schedule(dynamic, 64)
  for (int is = 0; is < DIM; ++is) {
    R t10[10], uu[4], ee;
    for (int c = 0; c < 10; ++c) t10[c] = T[(size_t)c * DIM + is];
    landau_cell<R>(t10, P.TAU, &ee, uu);
    e[is] = ee;
    for (int c = 0; c < 4; ++c) u[(size_t)c * DIM + is] = uu[c];
  }
  for (int is = 0; is < DIM; ++is) p[is] = pressure_of<R>(e[is], P.EOS_TYPE);
  if (P.VISCOUS_MATCHING) viscous_stages<R>(DIM, TAU, T.data(), e.data(), u.data(), p.data(), Pi.data(), pi.data());
  double t3 = now();
  // Output pack src/FreestreamMilne.cpp:695-718.
  if (out17) {
    const double hb = (double)kHbarc;
    bool fill = (P.OUTPUTFORMAT == 2) || (P.OUTPUTFORMAT == 3);
    for (int is = 0; is < DIM; ++is) {
      double *o = out17;
      size_t D = (size_t)DIM;
      if (!fill) {
        for (int k = 0; k < 17; ++k) o[k * D + is] = 0.0;
        continue;
      }
      o[0 * D + is] = (double)e[is] * hb;
      o[1 * D + is] = (double)p[is] * hb;
      for (int c = 0; c < 4; ++c) o[(2 + c) * D + is] = (double)u[c * D + is];
      for (int c = 0; c < 10; ++c) o[(6 + c) * D + is] = (double)pi[c * D + is] * hb;
      o[16 * D + is] = (double)Pi[is] * hb;
    }
  }
  double t4 = now();
  if (T_out)
    for (size_t i = 0; i < (size_t)10 * DIM; ++i) T_out[i] = (double)T[i];
  if (extra) {
    // extra layout: e[DIM], u[4*DIM] (fm^-4), then baryon nB[DIM], V[4*DIM]
    size_t D = (size_t)DIM;
    for (size_t i = 0; i < D; ++i) extra[i] = (double)e[i];
    for (size_t i = 0; i < 4 * D; ++i) extra[D + i] = (double)u[i];
    if (!J.empty()) {
      // src/LandauMatch.cpp:370-392
      for (int is = 0; is < DIM; ++is) {
        R nb = (u[0 * D + is] * J[0 * D + is]) - (u[1 * D + is] * J[1 * D + is]) -
               (u[2 * D + is] * J[2 * D + is]) - (TAU * TAU * u[3 * D + is] * J[3 * D + is]);
        extra[5 * D + is] = (double)nb;
        for (int c = 0; c < 4; ++c)
          extra[6 * D + c * D + is] = (double)(J[c * D + is] - (nb * u[c * D + is]));
      }
    }
  }
  if (stage_seconds) {
    stage_seconds[0] = t1 - t0;  // convert + freeStream
    stage_seconds[1] = t2 - t1;  // hypertrig + stress tensor
    stage_seconds[2] = t3 - t2;  // eigen + EoS + Pi + pi^{mu nu}
    stage_seconds[3] = t4 - t3;  // output pack
  }
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

// Parameter handling, restated from src/FreestreamMilne.cpp:149-178, 316-332
// and src/FreeStream.cpp:389-417 (used by the tests to cross-check the
// library's own implementation).
void oracle_params_default(fsm_params *p) {
  memset(p, 0, sizeof(*p));
  p->OUTPUTFORMAT = 2;
  p->BARYON = 0;
  p->IC_ENERGY = 5;
  p->IC_BARYON = 1;
  p->ETA_WIDTH = 0.5;
  p->ETA_FLAT = 0.5;
  p->SIGMA = 1.0;
  p->SIGMA_B = 1.0;
  p->DIM_X = 101;
  p->DIM_Y = 101;
  p->DIM_ETA = 1;
  p->DIM_RAP = 1;
  p->DIM_PHIP = 51;
  p->DX = 0.1;
  p->DY = 0.1;
  p->DETA = 0.1;
  p->DRAP = 0.2;
  p->DTAU = 0.5;
  p->TAU0 = 0.1;
  p->EOS_TYPE = 1;
  p->E_FREEZE = 1.7;
  p->VISCOUS_MATCHING = 1;
  p->E_DEP_FS = 0;
  p->DIM = p->DIM_X * p->DIM_Y * p->DIM_ETA;
  p->TAU = p->TAU0 + p->DTAU;
}

// Returns the finalized parameters (derived fields + overrides).
int oracle_params_finalize(fsm_params *p, const float *e_in, long n) {
  p->DIM = p->DIM_X * p->DIM_Y * p->DIM_ETA;
  p->TAU = p->TAU0 + p->DTAU;
  if (e_in && n != p->DIM) return FSM_ERR_SIZE;
  if (p->E_DEP_FS == 1) {
    if (!e_in) return FSM_ERR_ARG;
    // getEnergyDependentTau on the converted profile, serial float sums.
    float numerator = 0.0, denominator = 0.0;
    for (int is = 0; is < p->DIM; ++is) {
      float eps = p->IC_ENERGY == 5 ? e_in[is] * (float)1.0 / (float)kHbarc + (float)1.0e-7 : e_in[is];
      numerator += (eps * eps);
      denominator += eps;
    }
    float e_T = numerator / (denominator + 1.0e-7);
    if (!mut(MUT_EDEP_NO_HBARC)) e_T = e_T * kHbarc;
    float tau_fs = p->TAU_R * pow((e_T / p->E_R), p->ALPHA);
    p->DTAU = tau_fs;
    p->TAU = p->TAU0 + p->DTAU;
  }
  float min_dx_dy = std::min(p->DX, p->DY);
  int nphip = int(ceil((2.0 * M_PI * p->DTAU) / min_dx_dy));
  p->DIM_PHIP = std::max(p->DIM_PHIP, nphip);
  return 0;
}

// calculatePressure (src/EquationOfState.cpp:5-80) on n energy densities
// (fm^-4): precision 0 = float (the reference's fmaf chain), 1 = double.
void oracle_pressure(int eos, int precision, long n, const double *e, double *p) {
  for (long i = 0; i < n; ++i)
    p[i] = precision == 0 ? (double)pressure_of<float>((float)e[i], eos) : pressure_of<double>(e[i], eos);
}

// The reference float path's calculateBulkPressure and
// calculateShearViscTensor on given float planes (T 10, e, u 4, p): Pi, pi 10.
void oracle_viscous_float(const fsm_params *P, const float *T, const float *e, const float *u, const float *p,
                          float *Pi, float *pi) {
  viscous_stages<float>(P->DIM, P->TAU, T, e, u, p, Pi, pi);
}

// Rapidity rule of the 3-D path: rule 0 (the reference's tan grid), or 1
// with the n Gauss-Hermite nodes x and weights w for the weight exp(-x^2).
int oracle_set_rapidity_rule(int rule, int n, const double *x, const double *w) {
  if (rule == 0) {
    g_rap_rule = 0;
    return 0;
  }
  if (rule != 1 || n < 1 || !x || !w) return -1;
  g_gh_x.assign(x, x + n);
  g_gh_w.assign(w, w + n);
  g_rap_rule = 1;
  return 0;
}

// Plants mistake `m` (enum Mutation) for the tests of the pins; 0 restores
// the oracle.  Returns the previous value.
int oracle_set_mutation(int m) {
  const int old = g_mutation;
  g_mutation = m;
  return old;
}

// Whole step.  precision: 0 = reference float path, 1 = FP64 oracle.
// e_in: DIM floats (GeV/fm^3).  out17: 17*DIM doubles (may be NULL).
// T_out: 10*DIM doubles in fm^-4 (may be NULL).  extra: 10*DIM doubles
// (e, u^mu in fm^-4, then nB, V^mu when BARYON) (may be NULL).
// stage_seconds: 4 doubles (may be NULL).  p must be finalized.
int oracle_run(const fsm_params *p, const float *e_in, const float *nB_in, int precision,
               double *out17, double *T_out, double *extra, double *stage_seconds) {
  if (!p || !e_in) return FSM_ERR_ARG;
  if (precision == 0) return run_pipeline<float>(*p, e_in, nB_in, out17, T_out, extra, stage_seconds);
  return run_pipeline<double>(*p, e_in, nB_in, out17, T_out, extra, stage_seconds);
}

// Natural bicubic spline pieces for the tests.
// z: nx*ny x-major.  Outputs zx, zy, zxy (nx*ny each).
void oracle_spline_derivs(int nx, int ny, const double *x, const double *y, const double *z,
                          double *zx, double *zy, double *zxy) {
  Spline2D s;
  s.init(nx, ny, x, y, z);
  memcpy(zx, s.zx.data(), sizeof(double) * nx * ny);
  memcpy(zy, s.zy.data(), sizeof(double) * nx * ny);
  memcpy(zxy, s.zxy.data(), sizeof(double) * nx * ny);
}

void oracle_spline_eval(int nx, int ny, const double *x, const double *y, const double *z, int npts,
                        const double *px, const double *py, double *out) {
  Spline2D s;
  s.init(nx, ny, x, y, z);
  for (int i = 0; i < npts; ++i) out[i] = s.eval(px[i], py[i]);
}

// Eigen-decomposition of one real 4x4 (row-major A): wr, wi, V (row-major,
// columns are eigenvectors), returns 1 on convergence.
int oracle_eig4(const double *A, double *wr, double *wi, double *V) {
  double M[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) M[i][j] = A[i * 4 + j];
  Eig4 r;
  eig4(M, &r);
  for (int k = 0; k < 4; ++k) {
    wr[k] = r.wr[k];
    wi[k] = r.wi[k];
  }
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 4; ++k) V[i * 4 + k] = r.v[i][k];
  return r.ok;
}

// Landau matching of a batch of cells in double: T[n][10] -> e[n], u[n][4].
void oracle_landau(long n, const double *T, float TAU, double *e, double *u) {
#pragma omp parallel for schedule(dynamic, 256)
  for (long i = 0; i < n; ++i) landau_cell<double>(T + 10 * i, TAU, e + i, u + 4 * i);
}

// Condition numbers of the Landau matching of each cell: the largest relative
// change of e (kappa_e) and of u^mu / u^tau (kappa_u) over `ntrial` random
// relative perturbations of size rel of the ten components, divided by rel.
// T: n x 10 doubles.
// normwise = 1: first-order worst case over every component moving by up to
// rel * T^{tau tau} (times 1/TAU per eta index) -- the size of the rounding
// error of a quadrature sum of positive samples -- from one-sided differences
// per component: kappa_e = sum_c |de/dT_c| s_c / e, kappa_u = max_mu
// sum_c |du^mu/dT_c| s_c / u^tau (in units of rel).  normwise = 0: `ntrial`
// random perturbations of each component relative to itself.
void oracle_landau_condition(long n, const double *T, float TAU, int ntrial, double rel, unsigned seed, int normwise,
                             double *kappa_e, double *kappa_u) {
  static const int eta_idx[10] = {0, 0, 0, 1, 0, 0, 1, 0, 1, 2};
#pragma omp parallel for schedule(dynamic, 256)
  for (long i = 0; i < n; ++i) {
    double e0, u0[4];
    landau_cell<double>(T + 10 * i, TAU, &e0, u0);
    double ke = 0.0, ku = 0.0;
    if (normwise) {
      double se = 0.0, su[4] = {0.0, 0.0, 0.0, 0.0};
      for (int c = 0; c < 10; ++c) {
        double tp[10];
        for (int k = 0; k < 10; ++k) tp[k] = T[10 * i + k];
        const double h = rel * fabs(T[10 * i]) * pow((double)TAU, -eta_idx[c]);
        tp[c] += h;
        double e1, u1[4];
        landau_cell<double>(tp, TAU, &e1, u1);
        se += fabs(e1 - e0);
        for (int m = 0; m < 4; ++m) su[m] += fabs(u1[m] - u0[m]);
      }
      ke = se / (fabs(e0) * rel);
      for (int m = 0; m < 4; ++m) ku = std::max(ku, su[m] / (fabs(u0[0]) * rel));
    } else {
      unsigned s = seed ^ (unsigned)(i * 2654435761u);
      for (int t = 0; t < ntrial; ++t) {
        double tp[10];
        for (int c = 0; c < 10; ++c) {
          s = s * 1664525u + 1013904223u;
          double r = ((s >> 8) * (1.0 / 16777216.0)) * 2.0 - 1.0;  // [-1, 1)
          tp[c] = T[10 * i + c] * (1.0 + rel * r);
        }
        double e1, u1[4];
        landau_cell<double>(tp, TAU, &e1, u1);
        ke = std::max(ke, fabs(e1 - e0) / (fabs(e0) * rel));
        for (int c = 0; c < 4; ++c) ku = std::max(ku, fabs(u1[c] - u0[c]) / (fabs(u0[0]) * rel));
      }
    }
    kappa_e[i] = ke;
    kappa_u[i] = ku;
  }
}

// Diagnostics of the reference on its float hydro fields (D cells; u 4 planes,
// pi 10 planes): out = 17 planes of D floats in the order R_Pi_Inv,
// R_pimunu_Inv, pi_dot_u[4], pi_mu_mu, beta[4], omega[6] (zero where the
// reference leaves its calloc'd arrays untouched).
//   calculateBulkInvReynolds    src/HydroValidity.cpp:8-20
//   calculateShearInvReynolds   src/HydroValidity.cpp:22-39
//   calculate_pi_dot_u          src/HydroValidity.cpp:41-77
//   calculate_pi_mu_mu          src/HydroValidity.cpp:79-101
//   calculateThermalVorticityTensor src/LandauMatch.cpp:395-477 (2-D index
//                               DIM_Y*ix + iy over the transverse interior)
//   temperatureFromEnergyDensity src/EquationOfState.cpp:83-94
static float oracle_temperature(float eps) {
  float a = 13.8997;
  float T = powf((eps / a), 0.25);
  return T;
}

void oracle_diagnostics(const fsm_params *p, const float *e, const float *u, const float *pr, const float *Pi,
                        const float *pi, float *out) {
  const long D = (long)p->DIM_X * p->DIM_Y * p->DIM_ETA;
  const float TAU = p->TAU;
  for (long k = 0; k < 17 * D; ++k) out[k] = 0.0f;
  float *R_Pi = out, *R_pi = out + D, *pdu = out + 2 * D, *pmm = out + 6 * D, *beta = out + 7 * D,
        *omega = out + 11 * D;
  for (long is = 0; is < D; ++is) {
    float pp;
    if (pr[is] < 1.0e-5) pp = 1.0e-5;
    else pp = pr[is];
    R_Pi[is] = Pi[is] / pp;
  }
  const float tau2 = TAU * TAU;
  for (long is = 0; is < D; ++is) {
    const float *s0 = pi;
    float num = s0[is] * s0[is] -
                2.0 * (s0[D + is] * s0[D + is] + s0[2 * D + is] * s0[2 * D + is] + tau2 * s0[3 * D + is] * s0[3 * D + is]) +
                (s0[4 * D + is] * s0[4 * D + is] + s0[7 * D + is] * s0[7 * D + is] +
                 tau2 * tau2 * s0[9 * D + is] * s0[9 * D + is]) +
                2.0 * (s0[5 * D + is] * s0[5 * D + is] + tau2 * s0[6 * D + is] * s0[6 * D + is] +
                       tau2 * s0[8 * D + is] * s0[8 * D + is]);
    float den;
    if (e[is] < 1.0e-5) den = 1.0e-5;
    else den = e[is] * e[is] + 3.0 * pr[is] * pr[is];
    R_pi[is] = sqrt(fabs(num)) / sqrt(fabs(den));
  }
  for (long is = 0; is < D; ++is) {
    float ut = u[is], ux = u[D + is], uy = u[2 * D + is], un = u[3 * D + is];
    ux = -ux;
    uy = -uy;
    un = -TAU * TAU * un;
    float pitt = pi[is], pitx = pi[D + is], pity = pi[2 * D + is], pitn = pi[3 * D + is], pixx = pi[4 * D + is],
          pixy = pi[5 * D + is], pixn = pi[6 * D + is], piyy = pi[7 * D + is], piyn = pi[8 * D + is],
          pinn = pi[9 * D + is];
    pdu[is] = pitt * ut + pitx * ux + pity * uy + pitn * un;
    pdu[D + is] = pitx * ut + pixx * ux + pixy * uy + pixn * un;
    pdu[2 * D + is] = pity * ut + pixy * ux + piyy * uy + piyn * un;
    pdu[3 * D + is] = pitn * ut + pixn * ux + piyn * uy + pinn * un;
    pixx = -pixx;
    piyy = -piyy;
    pinn = -TAU * TAU * pinn;
    pmm[is] = pitt + pixx + piyy + pinn;
  }
  const int NX = p->DIM_X, NY = p->DIM_Y;
  const float dx = p->DX, dy = p->DY;
  for (int ix = 1; ix < NX - 1; ++ix)
    for (int iy = 1; iy < NY - 1; ++iy) {
      const long is = (long)NY * ix + iy, px = (long)NY * (ix + 1) + iy, mx = (long)NY * (ix - 1) + iy,
                 py = (long)NY * ix + (iy + 1), my = (long)NY * ix + (iy - 1);
      float T = oracle_temperature(e[is]);
      beta[is] = u[is] / T;
      beta[D + is] = -u[D + is] / T;
      beta[2 * D + is] = -u[2 * D + is] / T;
      beta[3 * D + is] = 0.0f / T;
      float bxpy = -u[D + py] / oracle_temperature(e[py]), bxmy = -u[D + my] / oracle_temperature(e[my]);
      float bypx = -u[2 * D + px] / oracle_temperature(e[px]), bymx = -u[2 * D + mx] / oracle_temperature(e[mx]);
      omega[3 * D + is] = 0.5 * ((bxpy - bxmy) / 2.0 / dy - (bypx - bymx) / 2.0 / dx);
    }
}

// Serial float sums of src/FreestreamMilne.cpp:531-545 (T00, e in fm^-4).
void oracle_energy_sums(const fsm_params *p, const float *T00, const float *e, float *tot, float *inside) {
  const long D = (long)p->DIM_X * p->DIM_Y * p->DIM_ETA;
  float a = 0.0, b = 0.0;
  for (long is = 0; is < D; ++is) a += T00[is];
  if (p->DIM_ETA > 1) a *= (p->TAU * p->DX * p->DY * p->DETA);
  else a *= (p->TAU * p->DX * p->DY);
  for (long is = 0; is < D; ++is)
    if ((e[is] * kHbarc) > p->E_FREEZE) b += T00[is];
  if (p->DIM_ETA > 1) b *= (p->TAU * p->DX * p->DY * p->DETA);
  else b *= (p->TAU * p->DX * p->DY);
  *tot = a;
  *inside = b;
}

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
