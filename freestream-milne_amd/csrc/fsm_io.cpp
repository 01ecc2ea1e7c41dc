// fsm_io.cpp -- the host-side input adapters and file formats of the
// reference: its initial-condition generators (src/InitialConditions.cpp),
// its profile readers (src/FileIO.cpp:133-164, src/InitialConditions.cpp:127-355),
// its OUTPUTFORMAT=1 text writers (src/FileIO.cpp:11-131) and the energy
// sums it prints and writes after streaming (src/FreestreamMilne.cpp:531-559).
//
// Float/double mixing follows the reference expressions (this file is built
// without FMA contraction), so generated profiles and written files match the
// reference's bytes on the same inputs.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <fstream>
#include <mutex>
#include <string>

#include "../../include/fsm.h"

namespace {

// ---- rand() of the reference build ------------------------------------------
// The fluctuating generators call the C library's rand() without seeding it
// (src/InitialConditions.cpp:100,120,268).  With glibc that is the additive
// feedback generator of random_r (TYPE_3, degree 31, separation 3) seeded
// with 1: r[0] = seed; r[i] = 16807 r[i-1] mod (2^31 - 1) for 0 < i < 31;
// r[i] = r[i-31] for 31 <= i < 34; then r[i] = r[i-31] + r[i-3] (mod 2^32),
// and the k-th output is r[k+344] >> 1.  RAND_MAX = 2^31 - 1.
struct LibcRand {
  uint32_t ring[34];
  int pos = 0;
  void seed(unsigned s) {
    int32_t r[34];
    r[0] = (int32_t)(s == 0 ? 1 : s);
    for (int i = 1; i < 31; ++i) {
      const int64_t v = (16807LL * (int64_t)r[i - 1]) % 2147483647LL;
      r[i] = (int32_t)(v < 0 ? v + 2147483647LL : v);
    }
    for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
    for (int i = 0; i < 34; ++i) ring[i] = (uint32_t)r[i];
    pos = 0;  // ring[pos] is r[i - 34] for the next i
    for (int k = 34; k < 344; ++k) step();
  }
  uint32_t step() {
    // r[i] = r[i-31] + r[i-3]; ring holds r[i-34 .. i-1] cyclically
    const uint32_t v = ring[(pos + 3) % 34] + ring[(pos + 31) % 34];
    ring[pos] = v;
    pos = (pos + 1) % 34;
    return v;
  }
  int next() { return (int)(step() >> 1); }
};

std::mutex g_rand_mu;
LibcRand g_rand;
bool g_rand_seeded = false;

int libc_rand() {
  if (!g_rand_seeded) {
    g_rand.seed(1);
    g_rand_seeded = true;
  }
  return g_rand.next();
}
const int kRandMax = 2147483647;

// Cell coordinates of the generators (src/InitialConditions.cpp:34-37): a
// float product minus a double half-width, stored as float.
inline float coord(int i, int n, float d) { return (float)((float)i * d - ((float)(n - 1)) / 2.0 * d); }

inline void cell_of(int is, const fsm_params &p, int *ix, int *iy, int *ie) {
  *ix = is / (p.DIM_Y * p.DIM_ETA);
  *iy = (is - (p.DIM_Y * p.DIM_ETA * *ix)) / p.DIM_ETA;
  *ie = is - (p.DIM_Y * p.DIM_ETA * *ix) - (p.DIM_ETA * *iy);
}

// initializeEllipticalGauss / initializeEllipticalMCGauss
// (src/InitialConditions.cpp:47-125): e0 = 500 fm^-4 times a Gaussian with
// variances bx, by, beta; the fluctuating form multiplies by rand()/RAND_MAX.
void elliptical_gauss(const fsm_params &p, float bx, float by, float beta, bool mc, float *d) {
  const float e0 = 500.0f;
  for (int is = 0; is < p.DIM; ++is) {
    int ix, iy, ie;
    cell_of(is, p, &ix, &iy, &ie);
    const float x = coord(ix, p.DIM_X, p.DX), y = coord(iy, p.DIM_Y, p.DY), eta = coord(ie, p.DIM_ETA, p.DETA);
    // left to right as written in the reference: e0 [* rand()/RAND_MAX] * exp * exp * exp
    if (mc) {
      const float f = (float)libc_rand() / (float)kRandMax;
      d[is] = (float)(e0 * f * exp(-(1.0 / bx) * (x * x)) * exp(-(1.0 / by) * (y * y)) *
                      exp(-(1.0 / beta) * (eta * eta)));
    } else {
      d[is] = (float)(e0 * exp(-(1.0 / bx) * (x * x)) * exp(-(1.0 / by) * (y * y)) * exp(-(1.0 / beta) * (eta * eta)));
    }
  }
}

// initialize2Gaussians (src/InitialConditions.cpp:238-272): its norm is 0, so
// every cell is 0, but it still draws one rand() per cell.
void two_gaussians(const fsm_params &p, float *d) {
  for (int is = 0; is < p.DIM; ++is) {
    (void)libc_rand();
    d[is] = 0.0f;
  }
}

// Longitudinal envelope of the block readers (src/InitialConditions.cpp:166-176):
// flat for |eta| < ETA_FLAT, Gaussian tails of width ETA_WIDTH, plus 1e-3.
void eta_envelope(const fsm_params &p, float *d) {
  const float lower_tolerance = 1.0e-3f;
  for (int is = 0; is < p.DIM; ++is) {
    int ix, iy, ie;
    cell_of(is, p, &ix, &iy, &ie);
    const float eta = coord(ie, p.DIM_ETA, p.DETA);
    const float a = fabsf(eta) - p.ETA_FLAT;
    float arg = (float)((-1.0) * a * a / (2.0 * p.ETA_WIDTH * p.ETA_WIDTH));
    arg = arg * ((float)a < 0.0f ? 0.0f : 1.0f);
    d[is] = (float)(d[is] * exp(arg) + lower_tolerance);
  }
}

int read_profile_impl(const fsm_params &p, int format, const char *path, float *d) {
  std::ifstream in(path);
  if (!in) return FSM_ERR_FILE;
  const int NX = p.DIM_X, NY = p.DIM_Y, NE = p.DIM_ETA;
  auto at = [&](int ix, int iy, int ie) { return (NY * NE) * ix + NE * iy + ie; };
  float v = 0.0f;
  switch (format) {
    case FSM_PROFILE_XYEV: {
      // readDensityFile (src/FileIO.cpp:133-164): "x y eta value" lines
      // placed at the rounded grid index; indices outside the grid are
      // skipped (the reference writes past its array there)
      const float xmin = (float)((-1.0) * ((float)(NX - 1) / 2.0) * p.DX);
      const float ymin = (float)((-1.0) * ((float)(NY - 1) / 2.0) * p.DY);
      const float etamin = (float)((-1.0) * ((float)(NE - 1) / 2.0) * p.DETA);
      float x, y, eta;
      while (in >> x >> y >> eta >> v) {
        const int ix = (int)roundf((x - xmin) / p.DX);
        const int iy = (int)roundf((y - ymin) / p.DY);
        const int ie = (int)roundf((eta - etamin) / p.DETA);
        if (ix < 0 || ix >= NX || iy < 0 || iy >= NY || ie < 0 || ie >= NE) continue;
        d[at(ix, iy, ie)] = v;
      }
      return FSM_OK;
    }
    case FSM_PROFILE_CPUVH: {
      // readEnergyDensityCPUVH (src/InitialConditions.cpp:314-355): "x y eta
      // value" in eta, y, x order; value + 1e-7
      float dummy;
      for (int ie = 0; ie < NE; ++ie)
        for (int iy = 0; iy < NY; ++iy)
          for (int ix = 0; ix < NX; ++ix) {
            if (!(in >> dummy >> dummy >> dummy >> v)) return FSM_OK;
            d[at(ix, iy, ie)] = v * 1.0f + 1.0e-7f;
          }
      return FSM_OK;
    }
    case FSM_PROFILE_SUPERMC_BLOCK:
    case FSM_PROFILE_TRENTO_BLOCK: {
      // readEnergyDensitySuperMCBlock (x rows of y values) and
      // readEnergyDensityTRENTOBlock (12 header lines, then y rows of x
      // values), src/InitialConditions.cpp:127-236: the transverse block is
      // copied to every eta and multiplied by the eta envelope
      const bool trento = format == FSM_PROFILE_TRENTO_BLOCK;
      if (trento) {
        std::string line;
        for (int l = 0; l < 12; ++l) std::getline(in, line);
      }
      for (int a = 0; a < (trento ? NY : NX); ++a)
        for (int b = 0; b < (trento ? NX : NY); ++b) {
          if (!(in >> v)) break;
          const int ix = trento ? b : a, iy = trento ? a : b;
          for (int ie = 0; ie < NE; ++ie) d[at(ix, iy, ie)] = v;
        }
      eta_envelope(p, d);
      return FSM_OK;
    }
    case FSM_PROFILE_TRENTO3D_BLOCK: {
      // readEnergyDensityTRENTO3DBlock (src/InitialConditions.cpp:274-312):
      // x, y, eta order, value + 1e-3
      for (int ix = 0; ix < NX; ++ix)
        for (int iy = 0; iy < NY; ++iy)
          for (int ie = 0; ie < NE; ++ie) {
            if (!(in >> v)) return FSM_OK;
            d[at(ix, iy, ie)] = v + 1.0e-3f;
          }
      return FSM_OK;
    }
    default:
      return FSM_ERR_ARG;
  }
}

std::string join(const char *dir, const char *file) {
  std::string s = dir ? dir : "";
  if (!s.empty() && s.back() != '/') s += '/';
  return s + file;
}

}  // namespace

extern "C" {

void fsm_ic_seed(unsigned seed) {
  std::lock_guard<std::mutex> g(g_rand_mu);
  g_rand.seed(seed);
  g_rand_seeded = true;
}

int fsm_ic_rand(void) {
  std::lock_guard<std::mutex> g(g_rand_mu);
  return libc_rand();
}

int fsm_read_profile(const fsm_params *p, int format, const char *path, float *out, size_t n) {
  if (!p || !path || !out) return FSM_ERR_ARG;
  if (n != (size_t)p->DIM_X * p->DIM_Y * p->DIM_ETA) return FSM_ERR_SIZE;
  fsm_params q = *p;
  q.DIM = (int)n;
  memset(out, 0, n * sizeof(float));
  return read_profile_impl(q, format, path, out);
}

// src/FreestreamMilne.cpp:231-298: the initial-condition dispatch.
int fsm_initial_condition(const fsm_params *p, int which, const char *profile_dir, float *out, size_t n) {
  if (!p || !out) return FSM_ERR_ARG;
  if (n != (size_t)p->DIM_X * p->DIM_Y * p->DIM_ETA) return FSM_ERR_SIZE;
  fsm_params q = *p;
  q.DIM = (int)n;
  memset(out, 0, n * sizeof(float));
  std::lock_guard<std::mutex> g(g_rand_mu);
  if (which == FSM_IC_WHICH_ENERGY) {
    switch (q.IC_ENERGY) {
      case 1: elliptical_gauss(q, 15.0f, 15.0f, 15.0f, false, out); return FSM_OK;
      case 2: elliptical_gauss(q, 15.0f, 15.0f, 15.0f, true, out); return FSM_OK;
      case 3: return read_profile_impl(q, FSM_PROFILE_XYEV, join(profile_dir, "e.dat").c_str(), out);
      case 4: return read_profile_impl(q, FSM_PROFILE_CPUVH, join(profile_dir, "e.dat").c_str(), out);
      case 6: two_gaussians(q, out); return FSM_OK;
      default: return FSM_ERR_IC;  // 5 is the caller's vector; others are invalid
    }
  }
  if (which == FSM_IC_WHICH_BARYON) {
    switch (q.IC_BARYON) {
      case 1: elliptical_gauss(q, 2.0f, 3.0f, 1.0f, false, out); return FSM_OK;
      case 2: elliptical_gauss(q, 2.0f, 3.0f, 1.0f, true, out); return FSM_OK;
      case 3: return read_profile_impl(q, FSM_PROFILE_XYEV, join(profile_dir, "nB.dat").c_str(), out);
      default: return FSM_ERR_IC;
    }
  }
  return FSM_ERR_ARG;
}

// writeScalarToFile / writeVectorToFile (src/FileIO.cpp:11-81): one
// "x y eta value" line per cell in cell order, coordinates snapped to the
// grid with roundf, default ostream formatting (6 significant digits).
int fsm_write_scalar(const fsm_params *p, const char *path, const float *var) {
  if (!p || !path || !var) return FSM_ERR_ARG;
  std::ofstream f(path);
  if (!f) return FSM_ERR_FILE;
  const int NX = p->DIM_X, NY = p->DIM_Y, NE = p->DIM_ETA;
  for (int ix = 0; ix < NX; ++ix)
    for (int iy = 0; iy < NY; ++iy)
      for (int ie = 0; ie < NE; ++ie) {
        float x = (float)((float)ix * p->DX - (((float)(NX - 1)) / 2.0 * p->DX));
        x = p->DX * roundf(x / p->DX);
        float y = (float)((float)iy * p->DY - (((float)(NY - 1)) / 2.0 * p->DY));
        y = p->DY * roundf(y / p->DY);
        float eta = (float)((float)ie * p->DETA - (((float)(NE - 1)) / 2.0 * p->DETA));
        eta = p->DETA * roundf(eta / p->DETA);
        const int is = (NY * NE) * ix + NE * iy + ie;
        f << x << " " << y << " " << eta << " " << var[is] << "\n";
      }
  return f.good() ? FSM_OK : FSM_ERR_FILE;
}

// writeScalarToFileProjection (src/FileIO.cpp:85-108): the eta = 0 slice
// ((DIM_ETA-1)/2) as DIM_Y rows of DIM_X values.
int fsm_write_projection(const fsm_params *p, const char *path, const float *var) {
  if (!p || !path || !var) return FSM_ERR_ARG;
  std::ofstream f(path);
  if (!f) return FSM_ERR_FILE;
  const int NX = p->DIM_X, NY = p->DIM_Y, NE = p->DIM_ETA;
  for (int iy = 0; iy < NY; ++iy) {
    for (int ix = 0; ix < NX; ++ix) {
      const int ie = (NE - 1) / 2;
      f << var[(NY * NE) * ix + NE * iy + ie] << " ";
    }
    f << "\n";
  }
  return f.good() ? FSM_OK : FSM_ERR_FILE;
}

// src/FreestreamMilne.cpp:531-559: total energy on the grid after streaming
// and the part in cells with e*hbarc > E_FREEZE, both as the reference's
// serial float sums of T^tautau times tau dx dy (deta).
int fsm_energy_sums(const fsm_params *p, const float *T00, const float *e, float *total, float *inside) {
  if (!p || !T00 || !e) return FSM_ERR_ARG;
  const float hbarc = 0.197326938f;
  const int DIM = p->DIM_X * p->DIM_Y * p->DIM_ETA;
  float tot = 0.0f, in = 0.0f;
  for (int is = 0; is < DIM; ++is) tot += T00[is];
  for (int is = 0; is < DIM; ++is)
    if ((e[is] * hbarc) > p->E_FREEZE) in += T00[is];
  if (p->DIM_ETA > 1) {
    tot *= (p->TAU * p->DX * p->DY * p->DETA);
    in *= (p->TAU * p->DX * p->DY * p->DETA);
  } else {
    tot *= (p->TAU * p->DX * p->DY);
    in *= (p->TAU * p->DX * p->DY);
  }
  if (total) *total = tot;
  if (inside) *inside = in;
  return FSM_OK;
}

// output/energy_inside_switch_surf.dat (src/FreestreamMilne.cpp:555-559).
int fsm_write_energy_inside(const char *path, float inside) {
  if (!path) return FSM_ERR_ARG;
  std::ofstream f(path);
  if (!f) return FSM_ERR_FILE;
  const float hbarc = 0.197326938f;
  f << inside * hbarc << std::endl;
  return f.good() ? FSM_OK : FSM_ERR_FILE;
}

}  // extern "C"

// Joined path helper for the class shim (not exported).
std::string fsm_io_join(const char *dir, const char *file) { return join(dir, file); }
