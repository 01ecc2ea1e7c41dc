/*
 * fsm.h -- C ABI of the MI355X free-streaming + Landau-matching library
 * (libfsm.so, built from freestream-milne_amd/csrc).
 *
 * The reference exposes this path as a source-level C++ class
 * (class FREESTREAMMILNE, src/FreestreamMilne.cpp:28-87) whose per-call
 * configuration is the positional file ./freestream_input
 * (src/FileIO.cpp:166-245) and whose results are 17 std::vector<double>
 * (src/FreestreamMilne.cpp:102-136, filled at :677-718).  This header is the
 * plain-pointer boundary that replaces it: every entry point takes POD
 * structs, raw pointers and sizes, and returns an int status.  The header-only
 * C++ class in freestream-milne_amd/csrc/FreestreamMilne.h rebuilds the
 * reference class on top of it.
 *
 * Units follow the reference: energy density in GeV/fm^3 at the boundary,
 * fm^-4 inside, lengths in fm, times in fm/c.
 *
 * Cell index (every per-cell array): is = (DIM_Y*DIM_ETA)*ix + DIM_ETA*iy + ieta
 * (src/FreeStream.cpp:159-161).  Stress-tensor component order:
 * tt, tx, ty, tn, xx, xy, xn, yy, yn, nn (src/LandauMatch.cpp:46-55).
 */
#ifndef FSM_H_
#define FSM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors struct parameters, src/Parameter.h:2-33, field for field and in
 * the same order, so a reference caller can memcpy its struct. */
typedef struct fsm_params {
  int OUTPUTFORMAT;
  int BARYON;
  int IC_ENERGY;
  int IC_BARYON;
  float ETA_WIDTH;
  float ETA_FLAT;
  float SIGMA;
  float SIGMA_B;
  int DIM_X;
  int DIM_Y;
  int DIM_ETA;
  int DIM_RAP;
  int DIM_PHIP;
  float DX;
  float DY;
  float DETA;
  float DRAP;
  float DTAU;
  float TAU0;
  int EOS_TYPE;
  float E_FREEZE;
  int VISCOUS_MATCHING;
  int E_DEP_FS;
  float E_R;
  float TAU_R;
  float ALPHA;
  /* derived: DIM = DIM_X*DIM_Y*DIM_ETA, TAU = TAU0 + DTAU */
  int DIM;
  float TAU;
} fsm_params;

/* Status codes.  The reference's conventions are exit(-1) on a vector-size
 * mismatch (src/FreestreamMilne.cpp:259-264), exit(1) on a missing profile
 * file (src/FileIO.cpp:150-154) and return 0 with no outputs on an invalid
 * initial condition (src/FreestreamMilne.cpp:277-281); the library returns a
 * code instead and the C++ shim maps it back. */
enum {
  FSM_OK = 0,
  FSM_ERR_SIZE = -1,        /* input vector size != DIM */
  FSM_ERR_IC = -2,          /* IC_ENERGY / IC_BARYON not supported on this path */
  FSM_ERR_FILE = -3,        /* parameter or profile file missing */
  FSM_ERR_OOM = -4,         /* device or host allocation failed */
  FSM_ERR_DEVICE = -5,      /* a HIP call failed */
  FSM_ERR_ARG = -6,         /* invalid argument (NULL pointer, bad size) */
  FSM_ERR_UNSUPPORTED = -7, /* parameter combination outside this library */
  FSM_ERR_BUSY = -8,        /* serving queue full / no slot free */
  FSM_ERR_CANCELLED = -9,   /* request cancelled before it ran */
  FSM_ERR_TOO_LARGE = -10,  /* event larger than the fixed workspace of fsm_run_device_noalloc */
  FSM_ERR_NO_PLAN = -11     /* fsm_run_device_noalloc: the event shape has no tables in the context yet */
};

/* Number and order of the OUTPUTFORMAT=2 result vectors:
 * e, p, ut, ux, uy, un, pitt, pitx, pity, pitn, pixx, pixy, pixn, piyy,
 * piyn, pinn, Pi  (src/FreestreamMilne.cpp:51-67, :700-716). */
#define FSM_NUM_OUTPUTS 17
/* Device output of a baryon event: the 17 planes, then n_B (fm^-3) and
 * V^tau, V^x, V^y, V^eta -- 22 planes of DIM doubles. */
#define FSM_NUM_OUTPUTS_BARYON 22

/* Library version string. */
const char *fsm_version(void);
const char *fsm_status_string(int status);

/* Built-in defaults: src/FreestreamMilne.cpp:149-171 (E_R, TAU_R, ALPHA have
 * no reference default; they are set to 0 here).  Also fills DIM and TAU. */
void fsm_params_default(fsm_params *p);

/* Positional reader of a freestream_input file: src/FileIO.cpp:166-245.
 * Reads 26 "NAME value" lines in the fixed reference order, ignoring the
 * names.  Returns FSM_ERR_FILE (and leaves *p untouched) if the file cannot
 * be opened -- the reference prints a message and keeps its defaults. */
int fsm_params_read(fsm_params *p, const char *path);

/* Derived fields and run-time overrides, in the reference order:
 *   DIM, TAU                       src/FreestreamMilne.cpp:177-178
 *   E_DEP_FS: DTAU = TAU_R*(e_T/E_R)^ALPHA from the converted input
 *                                  src/FreestreamMilne.cpp:316-323,
 *                                  src/FreeStream.cpp:389-417
 *   DIM_PHIP = max(DIM_PHIP, ceil(2 pi DTAU / min(DX,DY)))
 *                                  src/FreestreamMilne.cpp:329-332
 * e_GeV (length n, may be NULL when E_DEP_FS == 0) is the input vector in
 * GeV/fm^3.  Returns FSM_ERR_SIZE if e_GeV != NULL and n != DIM. */
int fsm_params_finalize(fsm_params *p, const float *e_GeV, size_t n);

/* ------------------------------------------------------------------ */
/* Context: one per host thread; owns device workspace and event tables. */
typedef struct fsm_ctx fsm_ctx;

/* Creates a context on HIP device `device`. */
int fsm_create(int device, fsm_ctx **out);
void fsm_destroy(fsm_ctx *ctx);

/* Quadrature path selector for the streaming stage. */
enum {
  FSM_PATH_FP32 = 0,  /* FP32 VALU interpolation and quadrature; FP64 Landau */
  FSM_PATH_MIXED = 1  /* 2-D: interpolation + quadrature as one contraction
                         on FP16 MFMA (knots and weights split hi+lo, FP32
                         accumulation), FP32 sample loop where the spline can
                         undershoot; the same FP64 Landau frame as the FP32
                         path.  No counterpart in the reference. */
};
int fsm_set_path(fsm_ctx *ctx, int path);

/* Rapidity quadrature of 3-D events (no counterpart in the reference).
 * FSM_RAPIDITY_TAN is the reference's rule: nodes y = eta + tan(pi/2 w) on a
 * uniform w grid, weights the Jacobian dy/dw (src/FreeStream.cpp:174-175,
 * :340-345; src/LandauMatch.cpp:93-95).  FSM_RAPIDITY_GAUSS_HERMITE replaces
 * them by the DIM_RAP-point Gauss-Hermite rule for the Gaussian rapidity
 * profile of convertInitialDensity: nodes y = eta + sqrt(2) SIGMA x_k and
 * weights sqrt(2) SIGMA W_k (x_k, W_k for the weight exp(-x^2), as in the
 * reference's quadrature_tables/gaussHermite*.dat), the profile's
 * exp(-(eta - y)^2 / (2 SIGMA^2)) moving from the density into the weights
 * (the density keeps norm * cosh^2(eta - y)).  Applies to events with
 * DIM_ETA > 1.  The charge pass of BARYON = 1 streams with the same nodes (in
 * the reference it shares the energy pass's geometry, src/FreestreamMilne.cpp:
 * 386) and keeps its density on the uniform DRAP grid and its weight DRAP
 * (src/FreeStream.cpp:353-387, src/LandauMatch.cpp:110-143), as the reference
 * does under the tan grid.  Context-wide; plans are kept per rule. */
enum { FSM_RAPIDITY_TAN = 0, FSM_RAPIDITY_GAUSS_HERMITE = 1 };
int fsm_set_rapidity_rule(fsm_ctx *ctx, int rule);

/* Nodes x[0..n) (ascending) and weights w[0..n) of the n-point Gauss-Hermite
 * rule for the weight exp(-x^2) (no counterpart in the reference, which only
 * ships the 21-, 31- and 41-point tables as data files).  n in [1, 200]. */
int fsm_gauss_hermite(int n, double *x, double *w);

/* What the planner chose for an event shape (no counterpart in the
 * reference; for tuning reports).  Plans are cached per parameter set. */
typedef struct fsm_plan_info {
  int dim3;             /* 1 for a 3-D (DIM_ETA > 1) plan */
  int variant;          /* 2-D tile shape index (0 = widest) */
  int tile_w, tile_h;   /* cells per tile along x and y */
  int tiles;            /* tiles of one event */
  int n_phi;            /* azimuth samples */
  int n_offsets;        /* knot offsets of the MFMA form */
  int n_offset_groups;  /* LDS stages of the MFMA form */
  int n_vector_samples; /* samples kept on the FP32 loop in MFMA tiles */
  int lds_bytes;        /* dynamic LDS per workgroup on the context's path */
  int batch_variant;    /* tile shape of this event inside batches */
  int batch_tiles;      /* tiles of one event inside batches */
  int tuned;            /* 1 if batch_variant was chosen by the autotuner */
  float tune_us[3];     /* autotuner: stream time per event of each shape */
  int mfma_batch;       /* MFMA instructions (16x16x32) of one event's stream stage on the batch shape */
  int mfma_one;         /* the same on the one-event shape (both count every tile on the MFMA form) */
  double samples_kept;      /* (cell, rapidity, azimuth) samples the reference's masks keep */
  double samples_evaluated; /* samples the stream kernel evaluates (3-D: whole tiles per reached sample) */
  float est_batch_us;       /* estimated device time of the event inside a batch on the context's path
                               (the serving loop's launch cap uses it) */
} fsm_plan_info;
int fsm_plan_describe(fsm_ctx *ctx, const fsm_params *p, fsm_plan_info *info);

/* Tile shapes (no counterpart in the reference).  A one-event call uses the
 * latency shape (enough workgroups for one event to fill the GPU); a batch
 * uses the plan's batch shape, which the autotuner picks by timing every
 * shape on a synthetic batch of the event's size the first time the shape
 * enters a batch or the serving loop (on by default).  FSM_TILES_BATCH makes
 * one-event calls use the batch shape too, so that a result is bitwise the
 * same alone and batched. */
enum { FSM_TILES_AUTO = 0, FSM_TILES_BATCH = 1 };
int fsm_set_tile_policy(fsm_ctx *ctx, int policy);
int fsm_set_autotune(fsm_ctx *ctx, int on);
/* Runs the autotuner for p's shape on the context's path now; returns the
 * chosen batch shape in *variant (may be NULL). */
int fsm_autotune(fsm_ctx *ctx, const fsm_params *p, int *variant);

/* ------------------------------------------------------------------ */
/* Whole step with host buffers (the drop-in call).
 *
 * Replaces FREESTREAMMILNE::run_freestream_milne() with IC_ENERGY=5
 * (src/FreestreamMilne.cpp:139-741, vector ingest :251-271) followed by
 * output_to_vectors() (src/FreestreamMilne.cpp:102-136).
 *   p        in: parameters as read (defaults + file); out: finalized
 *            (DIM, TAU, DTAU and DIM_PHIP overrides applied).
 *   e_GeV    initial energy density, n floats, GeV/fm^3 (IC_ENERGY = 5,
 *            the reference's vector input); for any other IC_ENERGY the
 *            fm^-4 profile made by fsm_initial_condition / fsm_read_profile,
 *            used as is.
 *   out      FSM_NUM_OUTPUTS caller-owned arrays of DIM doubles each, in
 *            the order documented above.  When OUTPUTFORMAT is not 2 or 3
 *            they are zero-filled, as the reference leaves them
 *            (src/FreestreamMilne.cpp:695).
 *   tau_landau_match  out (may be NULL): TAU0 + DTAU after overrides
 *            (src/FreestreamMilne.cpp:326).
 * Synchronous; the library keeps no pointer to caller memory. */
int fsm_run(fsm_ctx *ctx, fsm_params *p, const float *e_GeV, size_t n,
            double *const out[FSM_NUM_OUTPUTS], float *tau_landau_match);

/* Same, plus the baryon extension (BARYON=1): nB_in is the initial baryon
 * density (n floats, fm^-3) fed through convertInitialChargeDensity
 * (src/FreeStream.cpp:353-387), a second freeStream pass
 * (src/FreestreamMilne.cpp:386) and calculateBaryonCurrent /
 * calculateBaryonDensity / calculateBaryonDiffusion
 * (src/LandauMatch.cpp:110-143, :370-392).  The reference writes these
 * results to files only (src/FreestreamMilne.cpp:666-673); here they come
 * back as nB_out (DIM doubles) and V_out[4] (DIM doubles each, V^tau, V^x,
 * V^y, V^eta).  Either output may be NULL. */
int fsm_run_baryon(fsm_ctx *ctx, fsm_params *p, const float *e_GeV,
                   const float *nB_in, size_t n,
                   double *const out[FSM_NUM_OUTPUTS], double *nB_out,
                   double *const V_out[4], float *tau_landau_match);
/* Device-resident form (p finalized, BARYON = 1): d_out holds
 * FSM_NUM_OUTPUTS_BARYON planes of DIM doubles.  In 2-D, DIM_RAP must be 1
 * (FSM_ERR_UNSUPPORTED otherwise).  With BARYON = 0 or d_nB = NULL this is
 * fsm_run_device. */
int fsm_run_baryon_device(fsm_ctx *ctx, const fsm_params *p, const float *d_e_GeV, const float *d_nB,
                          double *d_out, void *stream);

/* ------------------------------------------------------------------ */
/* Device-resident entry points.  All pointers are device pointers; `stream`
 * is a hipStream_t (NULL = default stream).  `p` must be finalized.  Calls
 * are asynchronous on `stream`; tables for the event shape are built on the
 * first call with that shape and cached in the context.
 *
 * Ordering: the one-event calls of a context share its workspace (T^{mu nu}
 * scratch, spline planes, 3-D source planes), so each call's stream first
 * waits for the previous call's work on that context, whatever stream it was
 * queued on, and concurrent host threads take turns.  Two calls on different
 * streams therefore do not overlap on the device; for overlapping events use
 * a batch, the server, or one context per stream.  A stream under graph
 * capture is not made to wait (the capturing caller orders its graph). */

/* Whole step: e_GeV[DIM] -> out[17*DIM] doubles (SoA, FSM_NUM_OUTPUTS planes
 * of DIM in the output order).  Equivalent to fsm_run without host copies. */
int fsm_run_device(fsm_ctx *ctx, const fsm_params *p, const float *d_e_GeV,
                   double *d_out, void *stream);

/* fsm_run_device without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and results as fsm_run_device; its event-sized workspace is a fixed
 * 80 MiB device array that every GPU the library runs on holds until the
 * process ends (2-D events up to about 1.1 M cells, 3-D events up to about
 * 1.9 M cells: the 131^2 x 101 event needs 72.7 MiB); a larger event returns
 * FSM_ERR_TOO_LARGE.  The event shape's tables must exist in the context
 * already (any earlier call with that shape builds them, e.g.
 * fsm_plan_describe or fsm_run_device), else FSM_ERR_NO_PLAN.  Every call on
 * one GPU shares that array: calls on one context are ordered as the other
 * one-event calls are, but calls through different contexts on the same GPU
 * must not overlap (they would race on the array). */
int fsm_run_device_noalloc(fsm_ctx *ctx, const fsm_params *p, const float *d_e_GeV, double *d_out, void *stream);

/* Stage A: convertInitialDensity + freeStream + calculateHypertrigTable +
 * calculateStressTensor fused (src/FreeStream.cpp:97-350,
 * src/LandauMatch.cpp:14-108).  Input e_GeV[DIM] in GeV/fm^3 (the ingest of
 * src/FreestreamMilne.cpp:267 is applied inside); output d_T[10*DIM]
 * floats, 10 planes of DIM in the reference layout (stressTensor[10][DIM]). */
int fsm_stream_tmunu_device(fsm_ctx *ctx, const fsm_params *p,
                            const float *d_e_GeV, float *d_T, void *stream);

/* Stage B: solveEigenSystem + calculatePressure + calculateBulkPressure +
 * calculateShearViscTensor (src/LandauMatch.cpp:145-367,
 * src/EquationOfState.cpp:5-80) on d_T[10*DIM] (reference layout, fm^-4).
 * Float outputs in fm^-4 (any may be NULL): d_e[DIM], d_u[4*DIM],
 * d_p[DIM], d_Pi[DIM], d_pi[10*DIM]. */
int fsm_landau_device(fsm_ctx *ctx, const fsm_params *p, const float *d_T,
                      float *d_e, float *d_u, float *d_p, float *d_Pi,
                      float *d_pi, void *stream);

/* ------------------------------------------------------------------ */
/* The reference's stage functions one by one, on its host layout: every
 * argument is an array of planes of DIM floats in host memory, one pointer
 * per component (the float** of src/Memoryf.cpp:5-13), fm^-4.  Each call
 * copies its inputs to the device, runs the stage on the context's path and
 * copies its outputs back before returning (synchronous).  p must have DIM
 * and TAU set (fsm_params_finalize).  freestream-milne_amd/csrc/fsm_stages.h
 * wraps them under the reference's own names and signatures. */

/* solveEigenSystem(float **stressTensor, float *energyDensity,
 * float **flowVelocity, parameters), src/LandauMatch.cpp:145-331: T (10
 * planes, component order tt, tx, ty, tn, xx, xy, xn, yy, yn, nn) -> e, u (4
 * planes), with the reference's selection rules and fallback. */
int fsm_solve_eigen_system(fsm_ctx *ctx, const fsm_params *p, const float *const T[10], float *e,
                           float *const u[4]);
/* calculatePressure(float *energyDensity, float *baryonDensity,
 * float *pressure, parameters), src/EquationOfState.cpp:5-80 (EOS_TYPE 1, 2,
 * 3); nB is not read by any of them and may be NULL. */
int fsm_calculate_pressure(fsm_ctx *ctx, const fsm_params *p, const float *e, const float *nB, float *pr);
/* calculateBulkPressure(float **stressTensor, float *energyDensity,
 * float *pressure, float *bulkPressure, parameters),
 * src/LandauMatch.cpp:332-344. */
int fsm_calculate_bulk_pressure(fsm_ctx *ctx, const fsm_params *p, const float *const T[10], const float *e,
                                const float *pr, float *Pi);
/* calculateShearViscTensor(float **stressTensor, float *energyDensity,
 * float **flowVelocity, float *pressure, float *bulkPressure,
 * float **shearTensor, parameters), src/LandauMatch.cpp:345-367. */
int fsm_calculate_shear_visc_tensor(fsm_ctx *ctx, const fsm_params *p, const float *const T[10], const float *e,
                                    const float *const u[4], const float *pr, const float *Pi, float *const pi[10]);

/* ------------------------------------------------------------------ */
/* Batches: many independent events (any mix of shapes, 2-D and 3-D) run with
 * one launch per stage over a work list of (event, tile) items, captured in a
 * HIP graph and replayed while the batch is unchanged.  No reference
 * counterpart: the reference runs one event per process call.
 *
 * fsm_batch_add takes finalized params and device buffers (input DIM floats
 * in GeV/fm^3, output 17*DIM doubles, as fsm_run_device) and returns the
 * event's index (>= 0) or a negative status; the pointers must stay valid
 * until fsm_batch_clear.  Workspace is allocated by the batch on its first
 * run after a change; that rebuild (and fsm_batch_clear, and destroy) waits
 * for the batch's previous fsm_batch_run to finish on its stream. */
typedef struct fsm_batch fsm_batch;
int fsm_batch_create(fsm_ctx *ctx, fsm_batch **out);
void fsm_batch_destroy(fsm_batch *b);
int fsm_batch_add(fsm_batch *b, const fsm_params *p, const float *d_e_GeV, double *d_out);
/* A baryon event (p->BARYON = 1): d_nB is its initial n_B and d_out holds
 * FSM_NUM_OUTPUTS_BARYON planes (see fsm_run_baryon_device). */
int fsm_batch_add_baryon(fsm_batch *b, const fsm_params *p, const float *d_e_GeV, const float *d_nB,
                         double *d_out);
int fsm_batch_clear(fsm_batch *b);
int fsm_batch_size(const fsm_batch *b);
/* use_graph: 1 (default) capture and replay a HIP graph; 0 launch directly
 * (stage timing with fsm_enable_timing is recorded only in this mode). */
int fsm_batch_set_graph(fsm_batch *b, int use_graph);
int fsm_batch_run(fsm_batch *b, void *stream);

/* ------------------------------------------------------------------ */
/* Serving: a request loop with continuous batching.  No reference
 * counterpart (the reference runs one event per process call).
 *
 * A worker thread owned by the server launches whatever requests have been
 * admitted whenever a launch slot is free, as one batch (one launch per
 * stage); requests that arrive while earlier batches run are admitted as
 * request slots and HBM budget allow, and their uploads and launches overlap
 * the running work.  Inputs are copied to the device at submit; results stay
 * in HBM until fetched or released. */
typedef struct fsm_server fsm_server;
typedef struct fsm_server_config {
  size_t hbm_budget_bytes; /* device memory for admitted requests (0: 16 GiB) */
  int max_batch_events;    /* events per launch (0: 256) */
  int max_inflight;        /* launches in flight (0: 5); with two or more, all but the last two
                              (at least one) take only 2-D requests */
  int max_requests;        /* admitted + completed-unreleased requests (0: 4096) */
  float max_batch_us;      /* estimated device time per launch, microseconds (0: 1000); a
                              launch always takes at least one request */
} fsm_server_config;

int fsm_server_create(fsm_ctx *ctx, const fsm_server_config *cfg, fsm_server **out);
/* Waits for launched work, drops queued requests, frees everything. */
void fsm_server_destroy(fsm_server *s);
/* Admits one event: p as read (it is finalized with the input, as fsm_run
 * does), e_GeV (n floats, GeV/fm^3) and, for BARYON = 1, nB (n floats;
 * NULL otherwise).  Returns a request id >= 0, FSM_ERR_BUSY when the request
 * slots or the HBM budget are exhausted (nothing admitted), or another
 * negative status. */
int64_t fsm_server_submit(fsm_server *s, const fsm_params *p, const float *e_GeV, const float *nB, size_t n);
/* Same with the inputs already in device memory (d_e_GeV, d_nB: n floats
 * each; d_nB only for BARYON = 1): the server copies them into its own
 * buffers on its upload stream, after the work already queued on `stream`
 * (a hipStream_t; NULL = default stream), so the caller may write its
 * buffers again once that copy is done (at the latest when the request is
 * reported by poll).  With E_DEP_FS = 1 the input is also copied to the host
 * for the reference's serial float sums (src/FreeStream.cpp:389-417).  No
 * reference counterpart: keeps event profiles in HBM end to end. */
int64_t fsm_server_submit_device(fsm_server *s, const fsm_params *p, const float *d_e_GeV, const float *d_nB,
                                 size_t n, void *stream);
/* Id of a completed request not yet reported by poll, or -1 if none completes
 * within timeout_us microseconds (0: do not wait). */
int64_t fsm_server_poll(fsm_server *s, int timeout_us);
/* Completed request: status of its run, submit-to-completion latency in
 * microseconds (may be NULL) and its device output (17*DIM doubles, 22*DIM
 * for a baryon request, may be
 * NULL), valid until released. */
int fsm_server_result(fsm_server *s, int64_t id, double *latency_us, double **d_out);
/* Copies a completed request's 17 output vectors (DIM doubles each) to host
 * arrays, then releases it. */
int fsm_server_fetch(fsm_server *s, int64_t id, double *const out[FSM_NUM_OUTPUTS]);
/* Same for a baryon request (its device output has FSM_NUM_OUTPUTS_BARYON
 * planes): also n_B and V^mu (either may be NULL). */
int fsm_server_fetch_baryon(fsm_server *s, int64_t id, double *const out[FSM_NUM_OUTPUTS], double *nB_out,
                            double *const V_out[4]);
/* Releases a completed request (its HBM returns to the budget). */
int fsm_server_release(fsm_server *s, int64_t id);
/* Cancels a request that has not been launched: FSM_OK (its result status is
 * FSM_ERR_CANCELLED), or FSM_ERR_BUSY if it is already running or done. */
int fsm_server_cancel(fsm_server *s, int64_t id);
/* Counters: requests admitted, completed, launches, bytes of HBM in use. */
int fsm_server_stats(fsm_server *s, int64_t *admitted, int64_t *completed, int64_t *launches,
                     size_t *bytes_in_use);

/* ------------------------------------------------------------------ */
/* Input adapters: the reference's initial conditions and
 * profile readers, as host functions that fill a DIM-float vector in cell
 * order.  A vector made here is the input of fsm_run / fsm_run_device with
 * the same params: for IC_ENERGY != 5 the library takes the input as the
 * fm^-4 density itself (no GeV conversion and no 1e-7 floor, as the
 * reference's dispatch, src/FreestreamMilne.cpp:231-281). */
enum { FSM_IC_WHICH_ENERGY = 0, FSM_IC_WHICH_BARYON = 1 };
/* The initial-condition dispatch, src/FreestreamMilne.cpp:231-298.
 * Energy (p->IC_ENERGY): 1 smooth elliptical Gaussian
 * (initializeEllipticalGauss(15, 15, 15), src/InitialConditions.cpp:47-72),
 * 2 its fluctuating form (initializeEllipticalMCGauss, :100-125), 3 the
 * "x y eta value" file <profile_dir>/e.dat (readDensityFile,
 * src/FileIO.cpp:133-164), 4 the CPU-VH file <profile_dir>/e.dat
 * (readEnergyDensityCPUVH, :314-355), 6 two fluctuating Gaussians
 * (initialize2Gaussians, :238-272; its norm is 0).  Baryon (p->IC_BARYON):
 * 1 initializeEllipticalGauss(2, 3, 1), 2 its fluctuating form, 3
 * <profile_dir>/nB.dat.  IC_ENERGY = 5 (the caller's vector) and unknown
 * codes return FSM_ERR_IC; a missing file FSM_ERR_FILE.  Units fm^-4
 * (energy) and fm^-3 (baryon). */
int fsm_initial_condition(const fsm_params *p, int which, const char *profile_dir, float *out, size_t n);
/* The fluctuating generators draw from the C library's rand() sequence of
 * the reference build (glibc, unseeded, i.e. seed 1), restated here so that a
 * fresh process reproduces the reference's profiles.  fsm_ic_seed restarts
 * it (srand semantics); fsm_ic_rand draws one value (0 .. 2^31-1). */
void fsm_ic_seed(unsigned seed);
int fsm_ic_rand(void);
/* Profile readers (src/FileIO.cpp:133-164, src/InitialConditions.cpp:127-355). */
enum {
  FSM_PROFILE_XYEV = 0,           /* readDensityFile: "x y eta value" lines */
  FSM_PROFILE_CPUVH = 1,          /* readEnergyDensityCPUVH: x y eta value in eta, y, x order, +1e-7 */
  FSM_PROFILE_SUPERMC_BLOCK = 2,  /* readEnergyDensitySuperMCBlock: x rows of y values, eta envelope */
  FSM_PROFILE_TRENTO_BLOCK = 3,   /* readEnergyDensityTRENTOBlock: 12 header lines, y rows of x values */
  FSM_PROFILE_TRENTO3D_BLOCK = 4  /* readEnergyDensityTRENTO3DBlock: x, y, eta order, +1e-3 */
};
int fsm_read_profile(const fsm_params *p, int format, const char *path, float *out, size_t n);

/* ------------------------------------------------------------------ */
/* OUTPUTFORMAT=1 text writers, byte-for-byte the
 * reference's: writeScalarToFile / writeVectorToFile (src/FileIO.cpp:11-81,
 * "x y eta value" per cell) and the eta = 0 projections
 * (writeScalarToFileProjection, :85-131).  `path` is the file to write (the
 * reference writes output/<name>.dat under the working directory). */
int fsm_write_scalar(const fsm_params *p, const char *path, const float *var);
int fsm_write_projection(const fsm_params *p, const char *path, const float *var);
/* Energy after streaming and inside the freeze-out surface
 * (src/FreestreamMilne.cpp:531-553), as the reference's serial float sums of
 * T^tautau (host arrays, p finalized); fsm_write_energy_inside writes
 * output/energy_inside_switch_surf.dat (:555-559). */
int fsm_energy_sums(const fsm_params *p, const float *T00, const float *e, float *total, float *inside);
int fsm_write_energy_inside(const char *path, float inside);

/* ------------------------------------------------------------------ */
/* Diagnostics on the device, from the float hydro
 * fields of fsm_landau_device (all device pointers, DIM floats per plane;
 * u has 4 planes, pi 10).  Outputs (any may be NULL):
 *   R_Pi_Inv, R_pimunu_Inv   calculateBulkInvReynolds / calculateShearInvReynolds
 *                            (src/HydroValidity.cpp:8-39)
 *   pi_dot_u[4], pi_mu_mu    calculate_pi_dot_u / calculate_pi_mu_mu (:41-101)
 *   beta[4], omega[6]        calculateThermalVorticityTensor
 *                            (src/LandauMatch.cpp:395-489, with the conformal
 *                            T(e) of src/EquationOfState.cpp:83-94); like the
 *                            reference it indexes cells as DIM_Y*ix + iy over
 *                            the interior of the transverse grid and leaves the
 *                            other entries 0. */
typedef struct fsm_diag_out {
  float *R_Pi_Inv;
  float *R_pimunu_Inv;
  float *pi_dot_u[4];
  float *pi_mu_mu;
  float *beta[4];
  float *omega[6];
} fsm_diag_out;
int fsm_diagnostics_device(fsm_ctx *ctx, const fsm_params *p, const float *d_e, const float *d_u, const float *d_p,
                           const float *d_Pi, const float *d_pi, const fsm_diag_out *out, void *stream);

/* ------------------------------------------------------------------ */
/* The whole reference run with its files: run_freestream_milne()
 * (src/FreestreamMilne.cpp:139-741) for one event from its initial density
 * (e_in as fsm_run takes it; nB_in for BARYON = 1, fm^-3): the step, the
 * diagnostics, the energy sums, the 17 result vectors (out, may be NULL; zero
 * unless OUTPUTFORMAT is 2 or 3) and, when output_dir names an existing
 * directory, every output/<name>.dat file the reference writes in the same
 * format.  rep (may be NULL) receives the values the reference prints. */
typedef struct fsm_run_report {
  float total_energy_before;  /* src/FreestreamMilne.cpp:350-354 */
  float total_energy_after;   /* :531-535 */
  float energy_inside;        /* :537-545 (fraction = energy_inside / total_energy_after) */
  float R_Pi_Inv_center;      /* :566-568 */
  float R_pimunu_Inv_center;
  float tau_landau_match;     /* :326 */
  float dtau;                 /* DTAU after the E_DEP_FS rule */
  int dim_phip;               /* DIM_PHIP after the arc-length rule */
} fsm_run_report;
int fsm_run_files(fsm_ctx *ctx, fsm_params *p, const float *e_in, const float *nB_in, size_t n,
                  double *const out[FSM_NUM_OUTPUTS], const char *output_dir, fsm_run_report *rep);

/* Waits for all work the context queued on `stream`. */
int fsm_sync(fsm_ctx *ctx, void *stream);

/* Kernel-time counters of the last device call on this context (HIP events
 * recorded around each kernel on the launch stream), in milliseconds.
 * which: 0 = spline precompute, 1 = stream + quadrature, 2 = Landau + output
 * pack.  Timing is recorded only after fsm_enable_timing(ctx, 1). */
int fsm_enable_timing(fsm_ctx *ctx, int on);
int fsm_last_kernel_ms(fsm_ctx *ctx, int which, float *ms);

/* Sums of the same per-kernel times over every device call made since timing
 * was enabled or last collected (ms[3], in milliseconds), and the number of
 * calls; resets the sums.  Waits for the recorded work to finish. */
int fsm_kernel_time_totals(fsm_ctx *ctx, double ms[3], int *ncalls);

#ifdef __cplusplus
}
#endif

#endif /* FSM_H_ */
