// run_wrapper.cpp -- command-line driver in the shape of the reference's
// RunWrapper (src/RunWrapper.cpp:3-74): reads ./freestream_input, feeds an
// initial energy-density vector through FREESTREAMMILNE and reports the
// results.
//
//   fsm_run [profile.txt [out.txt]]
//
// profile.txt holds DIM_X*DIM_Y*DIM_ETA whitespace-separated values in GeV/fm^3
// in cell order (ix slowest, ieta fastest).  Without it, a Gaussian of
// 15 GeV/fm^3 and width 1 fm is used.  out.txt, when given, receives one line
// per cell with the 17 output values in output_to_vectors order.
#include <math.h>
#include <stdio.h>

#include <vector>

#include "FreestreamMilne.h"

int main(int argc, char **argv) {
  fsm_params p;
  fsm_params_default(&p);
  fsm_params_read(&p, "freestream_input");
  p.DIM = p.DIM_X * p.DIM_Y * p.DIM_ETA;
  std::vector<float> e((size_t)p.DIM);
  if (argc > 1) {
    FILE *f = fopen(argv[1], "r");
    if (!f) {
      printf("Couldn't open initial profile!\n");
      return 1;
    }
    size_t n = 0;
    float v;
    while (n < e.size() && fscanf(f, "%f", &v) == 1) e[n++] = v;
    fclose(f);
    e.resize(n);
  } else {
    const float xmin = -0.5f * (p.DIM_X - 1) * p.DX, ymin = -0.5f * (p.DIM_Y - 1) * p.DY;
    for (int ix = 0; ix < p.DIM_X; ++ix)
      for (int iy = 0; iy < p.DIM_Y; ++iy)
        for (int ie = 0; ie < p.DIM_ETA; ++ie) {
          float x = ix * p.DX + xmin, y = iy * p.DY + ymin;
          e[((size_t)ix * p.DIM_Y + iy) * p.DIM_ETA + ie] = 15.0f * expf(-(x * x + y * y) / 2.0f);
        }
  }
  FREESTREAMMILNE fs;
  fs.initialize_from_vector(e);
  fs.run_freestream_milne();
  std::vector<double> out[FSM_NUM_OUTPUTS];
  fs.output_to_vectors(out[0], out[1], out[2], out[3], out[4], out[5], out[6], out[7], out[8], out[9], out[10],
                       out[11], out[12], out[13], out[14], out[15], out[16]);
  double emax = 0.0;
  for (double v : out[0]) emax = v > emax ? v : emax;
  printf("tau_LandauMatch = %f fm/c, cells = %zu, max e = %g GeV/fm^3\n", fs.tau_LandauMatch, out[0].size(), emax);
  if (argc > 2) {
    FILE *f = fopen(argv[2], "w");
    if (!f) return 1;
    for (size_t i = 0; i < out[0].size(); ++i) {
      for (int k = 0; k < FSM_NUM_OUTPUTS; ++k) fprintf(f, k ? " %.9g" : "%.9g", out[k][i]);
      fprintf(f, "\n");
    }
    fclose(f);
  }
  printf("run_freestream_milne() ran sucessfully \n");
  return 0;
}
