// FreestreamMilne.h -- header-only C++ rebuild of the reference's drop-in
// class FREESTREAMMILNE (src/FreestreamMilne.cpp:28-136) on top of the C ABI
// in include/fsm.h.  Same members, same method names and argument meaning:
//
//   FREESTREAMMILNE fs;
//   fs.initialize_from_vector(e);     // GeV/fm^3, DIM_X*DIM_Y*DIM_ETA floats
//   fs.run_freestream_milne();        // reads ./freestream_input positionally
//   fs.output_to_vectors(e, p, ut, ux, uy, un, pitt, ..., Pi);
//   float tau = fs.tau_LandauMatch;
//
// Behaviour kept from the reference: a missing ./freestream_input prints a
// message and uses the built-in defaults (src/FileIO.cpp:178-181); the
// initial condition is dispatched on IC_ENERGY / IC_BARYON as in
// src/FreestreamMilne.cpp:231-298 (5 = the vector given to
// initialize_from_vector, the others from the reference's generators and
// ./initial_profiles files); a vector whose size differs from DIM prints the
// two sizes and calls exit(-1) (:259-264); an unknown initial condition
// prints "Not a valid initial Condition... Goodbye" and returns 0 (:277-281);
// a missing profile file prints the reference's message and calls exit(1)
// (src/FileIO.cpp:150-154); the output/*.dat files are written into ./output
// when that directory exists, and the reference's summary lines are printed.
#ifndef FSM_FREESTREAMMILNE_H_
#define FSM_FREESTREAMMILNE_H_

#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "fsm.h"

class FREESTREAMMILNE {
 public:
  FREESTREAMMILNE(int device = 0) : gridSize(0), tau_LandauMatch(0.0f), ctx_(nullptr), device_(device) {}
  ~FREESTREAMMILNE() {
    if (ctx_) fsm_destroy(ctx_);
  }
  FREESTREAMMILNE(const FREESTREAMMILNE &) = delete;
  FREESTREAMMILNE &operator=(const FREESTREAMMILNE &) = delete;

  int run_freestream_milne() {
    fsm_params p;
    fsm_params_default(&p);
    if (fsm_params_read(&p, "freestream_input") == FSM_ERR_FILE)
      printf("Couldn't open parameters.dat . Using default values!\n");
    p.DIM = p.DIM_X * p.DIM_Y * p.DIM_ETA;
    std::vector<float> e_in, nb_in;
    if (p.IC_ENERGY == 5) {
      if ((size_t)p.DIM != init_energy_density.size()) {
        printf("Grid dimension of input vector does not match freestream_input file! \n");
        printf("Vector size : %d , DIM = %d \n", (int)init_energy_density.size(), p.DIM);
        exit(-1);
      }
      e_in = init_energy_density;
    } else if (!make_ic(p, FSM_IC_WHICH_ENERGY, e_in)) {
      return 0;
    }
    if (p.BARYON && !make_ic(p, FSM_IC_WHICH_BARYON, nb_in)) return 0;
    if (!ctx_) {
      int st = fsm_create(device_, &ctx_);
      if (st != FSM_OK) {
        printf("fsm_create: %s\n", fsm_status_string(st));
        exit(-1);
      }
    }
    std::vector<double> *outs[FSM_NUM_OUTPUTS] = {
        &final_energy_density, &final_pressure, &final_ut,   &final_ux,   &final_uy,   &final_un,
        &final_pitt,           &final_pitx,     &final_pity, &final_pitn, &final_pixx, &final_pixy,
        &final_pixn,           &final_piyy,     &final_piyn, &final_pinn, &final_Pi};
    double *ptrs[FSM_NUM_OUTPUTS];
    for (int k = 0; k < FSM_NUM_OUTPUTS; ++k) {
      outs[k]->assign((size_t)p.DIM, 0.0);
      ptrs[k] = outs[k]->data();
    }
    fsm_run_report rep;
    int st = fsm_run_files(ctx_, &p, e_in.data(), p.BARYON ? nb_in.data() : nullptr, e_in.size(), ptrs, "output",
                           &rep);
    if (st != FSM_OK) {
      printf("run_freestream_milne: %s\n", fsm_status_string(st));
      exit(-1);
    }
    tau_LandauMatch = rep.tau_landau_match;
    printf("Total energy on grid before streaming : %f \n", rep.total_energy_before);
    printf("Total energy on grid after streaming : %f \n", rep.total_energy_after);
    printf("Fraction of energy contained in Freezeout Hypersurface : %f \n",
           rep.energy_inside / rep.total_energy_after);
    printf("R_Pi_Inv at center : %f \n", rep.R_Pi_Inv_center);
    printf("R_pimunu_Inv at center : %f \n", rep.R_pimunu_Inv_center);
    return 0;
  }

  // the total number of grid points; unused by the reference as well
  int gridSize;
  // TAU0 + DTAU after the run-time overrides (src/FreestreamMilne.cpp:326)
  float tau_LandauMatch;

  // GeV/fm^3 (src/FreestreamMilne.cpp:97-99: stored by value)
  void initialize_from_vector(std::vector<float> energy_density_in) { init_energy_density = energy_density_in; }
  std::vector<float> init_energy_density;

  void output_to_vectors(std::vector<double> &energy_density_out, std::vector<double> &pressure_out,
                         std::vector<double> &ut_out, std::vector<double> &ux_out, std::vector<double> &uy_out,
                         std::vector<double> &un_out, std::vector<double> &pitt_out,
                         std::vector<double> &pitx_out, std::vector<double> &pity_out,
                         std::vector<double> &pitn_out, std::vector<double> &pixx_out,
                         std::vector<double> &pixy_out, std::vector<double> &pixn_out,
                         std::vector<double> &piyy_out, std::vector<double> &piyn_out,
                         std::vector<double> &pinn_out, std::vector<double> &Pi_out) {
    energy_density_out = final_energy_density;
    pressure_out = final_pressure;
    ut_out = final_ut;
    ux_out = final_ux;
    uy_out = final_uy;
    un_out = final_un;
    pitt_out = final_pitt;
    pitx_out = final_pitx;
    pity_out = final_pity;
    pitn_out = final_pitn;
    pixx_out = final_pixx;
    pixy_out = final_pixy;
    pixn_out = final_pixn;
    piyy_out = final_piyy;
    piyn_out = final_piyn;
    pinn_out = final_pinn;
    Pi_out = final_Pi;
  }

  std::vector<double> final_energy_density;
  std::vector<double> final_pressure;
  std::vector<double> final_ut;
  std::vector<double> final_ux;
  std::vector<double> final_uy;
  std::vector<double> final_un;
  std::vector<double> final_pitt;
  std::vector<double> final_pitx;
  std::vector<double> final_pity;
  std::vector<double> final_pitn;
  std::vector<double> final_pixx;
  std::vector<double> final_pixy;
  std::vector<double> final_pixn;
  std::vector<double> final_piyy;
  std::vector<double> final_piyn;
  std::vector<double> final_pinn;
  std::vector<double> final_Pi;

 private:
  // src/FreestreamMilne.cpp:231-298 for one density; false (after the
  // reference's message) when the code is not a valid initial condition
  bool make_ic(fsm_params &p, int which, std::vector<float> &v) {
    v.assign((size_t)p.DIM, 0.0f);
    const int st = fsm_initial_condition(&p, which, "initial_profiles", v.data(), v.size());
    if (st == FSM_ERR_FILE) {
      printf("Couldn't open initial profile!\n");
      exit(1);
    }
    if (st != FSM_OK) {
      printf("Not a valid initial Condition... Goodbye\n");
      return false;
    }
    return true;
  }

  fsm_ctx *ctx_;
  int device_;
};

#endif  // FSM_FREESTREAMMILNE_H_
