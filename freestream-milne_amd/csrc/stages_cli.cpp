// stages_cli.cpp -- the reference's Landau-matching sequence
// (src/FreestreamMilne.cpp:443-506) written against fsm_stages.h, as a
// program built from the reference's orchestrator calls it: the stage
// functions by their own names on float** arrays allocated the reference's
// way (one calloc per component, src/Memoryf.cpp:5-13).
//
//   fsm_stages in.bin out.bin
//
// in.bin: an fsm_params (DIM and TAU set) followed by the stress tensor as 10
// planes of DIM floats (fm^-4).  out.bin: e, u (4 planes), p, Pi, pi (10
// planes), DIM floats each.
#include <stdio.h>
#include <stdlib.h>

#include "fsm_stages.h"

static float **calloc2d(int n, int dim) {
  float **a = (float **)calloc(n, sizeof(float *));
  for (int i = 0; i < n; ++i) a[i] = (float *)calloc(dim, sizeof(float));
  return a;
}

static void free2d(float **a, int n) {
  for (int i = 0; i < n; ++i) free(a[i]);
  free(a);
}

int main(int argc, char **argv) {
  if (argc < 3) {
    printf("usage: fsm_stages in.bin out.bin\n");
    return 2;
  }
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 1;
  parameters params;
  if (fread(&params, sizeof(params), 1, f) != 1 || params.DIM <= 0) return 1;
  const int DIM = params.DIM;
  float **stressTensor = calloc2d(10, DIM);
  for (int c = 0; c < 10; ++c)
    if (fread(stressTensor[c], sizeof(float), DIM, f) != (size_t)DIM) return 1;
  fclose(f);
  float *energyDensity = (float *)calloc(DIM, sizeof(float));
  float *baryonDensity = params.BARYON ? (float *)calloc(DIM, sizeof(float)) : NULL;
  float **flowVelocity = calloc2d(4, DIM);
  float *pressure = (float *)calloc(DIM, sizeof(float));
  float *bulkPressure = (float *)calloc(DIM, sizeof(float));
  float **shearTensor = calloc2d(10, DIM);

  solveEigenSystem(stressTensor, energyDensity, flowVelocity, params);
  calculatePressure(energyDensity, baryonDensity, pressure, params);
  if (params.VISCOUS_MATCHING) {
    calculateBulkPressure(stressTensor, energyDensity, pressure, bulkPressure, params);
    calculateShearViscTensor(stressTensor, energyDensity, flowVelocity, pressure, bulkPressure, shearTensor, params);
  }

  FILE *o = fopen(argv[2], "wb");
  if (!o) return 1;
  fwrite(energyDensity, sizeof(float), DIM, o);
  for (int c = 0; c < 4; ++c) fwrite(flowVelocity[c], sizeof(float), DIM, o);
  fwrite(pressure, sizeof(float), DIM, o);
  fwrite(bulkPressure, sizeof(float), DIM, o);
  for (int c = 0; c < 10; ++c) fwrite(shearTensor[c], sizeof(float), DIM, o);
  fclose(o);
  free2d(stressTensor, 10);
  free2d(flowVelocity, 4);
  free2d(shearTensor, 10);
  free(energyDensity);
  free(baryonDensity);
  free(pressure);
  free(bulkPressure);
  printf("stages done, DIM = %d\n", DIM);
  return 0;
}
