// fsm_stages.h -- the reference's Landau-matching stage functions under their
// own names and signatures, run on the GPU through the C ABI in
// include/fsm.h.  A program built from the reference's sources swaps its CPU
// stages for these by compiling without src/LandauMatch.cpp:145-367 and
// src/EquationOfState.cpp:5-80 and including this header instead:
//
//   solveEigenSystem(stressTensor, energyDensity, flowVelocity, params);
//   calculatePressure(energyDensity, baryonDensity, pressure, params);
//   calculateBulkPressure(stressTensor, energyDensity, pressure, bulkPressure, params);
//   calculateShearViscTensor(stressTensor, energyDensity, flowVelocity, pressure, bulkPressure,
//                            shearTensor, params);
//
// The arrays are the reference's own (float** = one pointer per component,
// each DIM floats, src/Memoryf.cpp:5-13).  `parameters` is struct parameters
// of src/Parameter.h:2-33; fsm_params has the same fields in the same order.
// Define FSM_STAGES_REFERENCE_PARAMETERS before including this header when
// the reference's Parameter.h is included too; otherwise `parameters` is
// fsm_params.  Each thread uses its own library context on device
// $FSM_DEVICE (default 0), created on first use.  The reference's functions
// have no error return: a failure prints the library's status and calls
// exit(-1), as the reference does on its one error (src/FreestreamMilne.cpp:259-264).
#ifndef FSM_STAGES_H_
#define FSM_STAGES_H_

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fsm.h"

#ifndef FSM_STAGES_REFERENCE_PARAMETERS
typedef fsm_params parameters;
#endif

namespace fsm_stages_detail {

struct ThreadContext {
  fsm_ctx *ctx = nullptr;
  ~ThreadContext() {
    if (ctx) fsm_destroy(ctx);
  }
};

inline fsm_ctx *context() {
  thread_local ThreadContext t;
  if (!t.ctx) {
    const char *d = getenv("FSM_DEVICE");
    const int st = fsm_create(d ? atoi(d) : 0, &t.ctx);
    if (st != FSM_OK) {
      printf("fsm_create: %s\n", fsm_status_string(st));
      exit(-1);
    }
  }
  return t.ctx;
}

inline fsm_params params_of(const parameters &in) {
  static_assert(sizeof(parameters) == sizeof(fsm_params), "struct parameters and fsm_params differ");
  fsm_params p;
  memcpy(&p, &in, sizeof(p));
  return p;
}

inline void check(int st, const char *what) {
  if (st != FSM_OK) {
    printf("%s: %s\n", what, fsm_status_string(st));
    exit(-1);
  }
}

}  // namespace fsm_stages_detail

// src/LandauMatch.cpp:145-331
inline void solveEigenSystem(float **stressTensor, float *energyDensity, float **flowVelocity, parameters params) {
  const fsm_params p = fsm_stages_detail::params_of(params);
  fsm_stages_detail::check(fsm_solve_eigen_system(fsm_stages_detail::context(), &p, stressTensor, energyDensity,
                                                  flowVelocity),
                           "solveEigenSystem");
}

// src/EquationOfState.cpp:5-80
inline void calculatePressure(float *energyDensity, float *baryonDensity, float *pressure, parameters params) {
  const fsm_params p = fsm_stages_detail::params_of(params);
  fsm_stages_detail::check(fsm_calculate_pressure(fsm_stages_detail::context(), &p, energyDensity, baryonDensity,
                                                  pressure),
                           "calculatePressure");
}

// src/LandauMatch.cpp:332-344
inline void calculateBulkPressure(float **stressTensor, float *energyDensity, float *pressure, float *bulkPressure,
                                  parameters params) {
  const fsm_params p = fsm_stages_detail::params_of(params);
  fsm_stages_detail::check(fsm_calculate_bulk_pressure(fsm_stages_detail::context(), &p, stressTensor,
                                                       energyDensity, pressure, bulkPressure),
                           "calculateBulkPressure");
}

// src/LandauMatch.cpp:345-367
inline void calculateShearViscTensor(float **stressTensor, float *energyDensity, float **flowVelocity,
                                     float *pressure, float *bulkPressure, float **shearTensor, parameters params) {
  const fsm_params p = fsm_stages_detail::params_of(params);
  fsm_stages_detail::check(fsm_calculate_shear_visc_tensor(fsm_stages_detail::context(), &p, stressTensor,
                                                           energyDensity, flowVelocity, pressure, bulkPressure,
                                                           shearTensor),
                           "calculateShearViscTensor");
}

#endif  // FSM_STAGES_H_
