// fsm_params.cpp -- parameter defaults, the positional freestream_input
// reader and the run-time overrides of the reference, as host C++.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>

#include "../../include/fsm.h"

extern "C" {

// src/FreestreamMilne.cpp:149-171 (+ derived fields :177-178).
void fsm_params_default(fsm_params *p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->OUTPUTFORMAT = 2;
  p->BARYON = 0;
  p->IC_ENERGY = 5;
  p->IC_BARYON = 1;
  p->ETA_WIDTH = 0.5f;
  p->ETA_FLAT = 0.5f;
  p->SIGMA = 1.0f;
  p->SIGMA_B = 1.0f;
  p->DIM_X = 101;
  p->DIM_Y = 101;
  p->DIM_ETA = 1;
  p->DIM_RAP = 1;
  p->DIM_PHIP = 51;
  p->DX = 0.1f;
  p->DY = 0.1f;
  p->DETA = 0.1f;
  p->DRAP = 0.2f;
  p->DTAU = 0.5f;
  p->TAU0 = 0.1f;
  p->EOS_TYPE = 1;
  p->E_FREEZE = 1.7f;
  p->VISCOUS_MATCHING = 1;
  p->E_DEP_FS = 0;
  p->DIM = p->DIM_X * p->DIM_Y * p->DIM_ETA;
  p->TAU = p->TAU0 + p->DTAU;
}

// src/FileIO.cpp:166-245: 26 "NAME<ws>value" lines in fixed order; the names
// are read and discarded.  A missing file leaves *p unchanged.
int fsm_params_read(fsm_params *p, const char *path) {
  if (!p || !path) return FSM_ERR_ARG;
  FILE *f = fopen(path, "r");
  if (!f) return FSM_ERR_FILE;
  char name[256];
  int iv = 0;
  float fv = 0.0f;
  // Each entry: 'i' = int, 'f' = float, in the order of src/FileIO.cpp:185-240.
  struct Entry {
    char kind;
    void *dst;
  } order[] = {
      {'i', &p->OUTPUTFORMAT}, {'i', &p->BARYON},   {'i', &p->IC_ENERGY}, {'i', &p->IC_BARYON},
      {'f', &p->ETA_WIDTH},    {'f', &p->ETA_FLAT}, {'f', &p->SIGMA},     {'f', &p->SIGMA_B},
      {'i', &p->DIM_X},        {'i', &p->DIM_Y},    {'i', &p->DIM_ETA},   {'i', &p->DIM_RAP},
      {'i', &p->DIM_PHIP},     {'f', &p->DX},       {'f', &p->DY},        {'f', &p->DETA},
      {'f', &p->DRAP},         {'f', &p->DTAU},     {'f', &p->TAU0},      {'i', &p->EOS_TYPE},
      {'f', &p->E_FREEZE},     {'i', &p->VISCOUS_MATCHING}, {'i', &p->E_DEP_FS},
      {'f', &p->E_R},          {'f', &p->TAU_R},    {'f', &p->ALPHA},
  };
  for (auto &e : order) {
    // fscanf(fileIn, "%s\t%d\n", ...): a failed conversion leaves the
    // previous dummy value, which the reference then assigns.
    if (e.kind == 'i') {
      if (fscanf(f, "%255s\t%d\n", name, &iv) < 0) {
      }
      *(int *)e.dst = iv;
    } else {
      if (fscanf(f, "%255s\t%f\n", name, &fv) < 0) {
      }
      *(float *)e.dst = fv;
    }
  }
  fclose(f);
  return FSM_OK;
}

// Derived fields and overrides: src/FreestreamMilne.cpp:177-178, 316-332;
// getEnergyDependentTau src/FreeStream.cpp:389-417 on the converted input.
int fsm_params_finalize(fsm_params *p, const float *e_GeV, size_t n) {
  if (!p) return FSM_ERR_ARG;
  p->DIM = p->DIM_X * p->DIM_Y * p->DIM_ETA;
  p->TAU = p->TAU0 + p->DTAU;
  if (e_GeV && n != (size_t)p->DIM) return FSM_ERR_SIZE;
  if (p->E_DEP_FS == 1) {
    if (!e_GeV) return FSM_ERR_ARG;
    const float hbarc = 0.197326938;
    float numerator = 0.0f, denominator = 0.0f;
    for (int is = 0; is < p->DIM; ++is) {
      // the converted vector (src/FreestreamMilne.cpp:267), or the fm^-4
      // profile of the other initial conditions as is
      float eps = p->IC_ENERGY == 5 ? e_GeV[is] * 1.0f / hbarc + 1.0e-7f : e_GeV[is];
      numerator += (eps * eps);
      denominator += eps;
    }
    float e_T = numerator / (denominator + 1.0e-7);
    e_T = e_T * hbarc;
    float tau_fs = p->TAU_R * powf((e_T / p->E_R), p->ALPHA);
    p->DTAU = tau_fs;
    p->TAU = p->TAU0 + p->DTAU;
  }
  float min_dx_dy = std::min(p->DX, p->DY);
  int nphip = int(ceil((2.0 * M_PI * p->DTAU) / min_dx_dy));
  p->DIM_PHIP = std::max(p->DIM_PHIP, nphip);
  return FSM_OK;
}

}  // extern "C"
