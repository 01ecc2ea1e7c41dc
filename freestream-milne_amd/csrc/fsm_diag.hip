// fsm_diag.hip -- the reference's hydro-validity and vorticity diagnostics
// on the device (HBM-bound, one thread per cell):
//   calculateBulkInvReynolds, calculateShearInvReynolds, calculate_pi_dot_u,
//   calculate_pi_mu_mu                        src/HydroValidity.cpp:8-101
//   calculateThermalVorticityTensor           src/LandauMatch.cpp:395-477
//   temperatureFromEnergyDensity (conformal)  src/EquationOfState.cpp:83-94
// Expressions keep the reference's float/double mixing; contraction is off.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include "../../include/fsm.h"

namespace {

struct DiagArgs {
  int dim, nx, ny;
  float tau, dx, dy;
  const float *e, *u, *p, *Pi, *pi;  // u: 4 planes, pi: 10 planes of dim
  fsm_diag_out out;
};

__device__ __forceinline__ float temperature_conformal(float eps) {
  // eps = a T^4 with a = 13.8997 (src/EquationOfState.cpp:83-94)
  const float a = 13.8997f;
  return powf((eps / a), 0.25f);
}

__global__ __launch_bounds__(256) void k_diagnostics(DiagArgs A) {
#pragma clang fp contract(off)
  const long D = A.dim;
  for (long is = (long)blockIdx.x * blockDim.x + threadIdx.x; is < D; is += (long)gridDim.x * blockDim.x) {
    const float *pi = A.pi;
    if (A.out.R_Pi_Inv) {
      // src/HydroValidity.cpp:8-20
      const float pv = A.p[is] < 1.0e-5f ? 1.0e-5f : A.p[is];
      A.out.R_Pi_Inv[is] = A.Pi[is] / pv;
    }
    if (A.out.R_pimunu_Inv) {
      // src/HydroValidity.cpp:22-39
      const float tau2 = A.tau * A.tau;
      float s[10];
#pragma unroll
      for (int c = 0; c < 10; ++c) s[c] = pi[c * D + is];
      const float num =
          (float)((double)(s[0] * s[0]) - 2.0 * (double)(s[1] * s[1] + s[2] * s[2] + tau2 * s[3] * s[3]) +
                  (double)(s[4] * s[4] + s[7] * s[7] + tau2 * tau2 * s[9] * s[9]) +
                  2.0 * (double)(s[5] * s[5] + tau2 * s[6] * s[6] + tau2 * s[8] * s[8]));
      const float e = A.e[is], p = A.p[is];
      float den;
      if (e < 1.0e-5f)
        den = 1.0e-5f;
      else
        den = (float)((double)(e * e) + 3.0 * (double)p * (double)p);
      A.out.R_pimunu_Inv[is] = sqrtf(fabsf(num)) / sqrtf(fabsf(den));
    }
    if (A.out.pi_dot_u[0] || A.out.pi_dot_u[1] || A.out.pi_dot_u[2] || A.out.pi_dot_u[3]) {
      // src/HydroValidity.cpp:41-77: pi^{mu nu} u_nu with u_nu = (u^t, -u^x, -u^y, -tau^2 u^eta)
      const float ut = A.u[is], ux = -A.u[D + is], uy = -A.u[2 * D + is], un = -A.tau * A.tau * A.u[3 * D + is];
      const float tt = pi[is], tx = pi[D + is], ty = pi[2 * D + is], tn = pi[3 * D + is], xx = pi[4 * D + is],
                  xy = pi[5 * D + is], xn = pi[6 * D + is], yy = pi[7 * D + is], yn = pi[8 * D + is],
                  nn = pi[9 * D + is];
      if (A.out.pi_dot_u[0]) A.out.pi_dot_u[0][is] = tt * ut + tx * ux + ty * uy + tn * un;
      if (A.out.pi_dot_u[1]) A.out.pi_dot_u[1][is] = tx * ut + xx * ux + xy * uy + xn * un;
      if (A.out.pi_dot_u[2]) A.out.pi_dot_u[2][is] = ty * ut + xy * ux + yy * uy + yn * un;
      if (A.out.pi_dot_u[3]) A.out.pi_dot_u[3][is] = tn * ut + xn * ux + yn * uy + nn * un;
    }
    if (A.out.pi_mu_mu) {
      // src/HydroValidity.cpp:79-101
      const float tt = pi[is], xx = -pi[4 * D + is], yy = -pi[7 * D + is], nn = -A.tau * A.tau * pi[9 * D + is];
      A.out.pi_mu_mu[is] = tt + xx + yy + nn;
    }
    // Thermal velocity and vorticity (src/LandauMatch.cpp:395-477).  The
    // reference loops over the transverse interior with the 2-D index
    // DIM_Y*ix + iy whatever DIM_ETA is; every other entry stays 0.
    bool any_b = false, any_w = false;
#pragma unroll
    for (int c = 0; c < 4; ++c) any_b |= A.out.beta[c] != nullptr;
#pragma unroll
    for (int c = 0; c < 6; ++c) any_w |= A.out.omega[c] != nullptr;
    if (any_b || any_w) {
      float b[4] = {0.0f, 0.0f, 0.0f, 0.0f}, w[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
      const int ix = (int)(is / A.ny), iy = (int)(is - (long)ix * A.ny);
      if (is < (long)A.nx * A.ny && ix >= 1 && ix < A.nx - 1 && iy >= 1 && iy < A.ny - 1) {
        const long px = (long)A.ny * (ix + 1) + iy, mx = (long)A.ny * (ix - 1) + iy;
        const long py = (long)A.ny * ix + (iy + 1), my = (long)A.ny * ix + (iy - 1);
        const float T = temperature_conformal(A.e[is]);
        b[0] = A.u[is] / T;
        b[1] = -A.u[D + is] / T;
        b[2] = -A.u[2 * D + is] / T;
        b[3] = 0.0f / T;
        const float bx_py = -A.u[D + py] / temperature_conformal(A.e[py]);
        const float bx_my = -A.u[D + my] / temperature_conformal(A.e[my]);
        const float by_px = -A.u[2 * D + px] / temperature_conformal(A.e[px]);
        const float by_mx = -A.u[2 * D + mx] / temperature_conformal(A.e[mx]);
        // w_xy = 1/2 (d_y beta_x - d_x beta_y), central differences
        w[3] = (float)(0.5 * ((double)(bx_py - bx_my) / 2.0 / (double)A.dy -
                              (double)(by_px - by_mx) / 2.0 / (double)A.dx));
      }
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (A.out.beta[c]) A.out.beta[c][is] = b[c];
#pragma unroll
      for (int c = 0; c < 6; ++c)
        if (A.out.omega[c]) A.out.omega[c][is] = w[c];
    }
  }
}

}  // namespace

hipError_t fsm_launch_diagnostics(const fsm_params &p, const float *e, const float *u, const float *pr,
                                  const float *Pi, const float *pi, const fsm_diag_out &out, hipStream_t s) {
  DiagArgs A;
  A.dim = p.DIM_X * p.DIM_Y * p.DIM_ETA;
  A.nx = p.DIM_X;
  A.ny = p.DIM_Y;
  A.tau = p.TAU;
  A.dx = p.DX;
  A.dy = p.DY;
  A.e = e;
  A.u = u;
  A.p = pr;
  A.Pi = Pi;
  A.pi = pi;
  A.out = out;
  const int blocks = (int)std::min<long>(((long)A.dim + 255) / 256, 65535L);
  hipLaunchKernelGGL(k_diagnostics, dim3(blocks), dim3(256), 0, s, A);
  return hipGetLastError();
}
