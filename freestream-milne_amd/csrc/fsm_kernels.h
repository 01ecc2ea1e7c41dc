// fsm_kernels.h -- host-side launchers of the kernels in fsm_kernels.hip.
#ifndef FSM_KERNELS_H_
#define FSM_KERNELS_H_

#include <hip/hip_runtime.h>

#include "../../include/fsm.h"
#include "fsm_internal.h"

hipError_t fsm_launch_spline2d(const Event2D &ev, hipStream_t s);
hipError_t fsm_launch_stream2d(const Event2D &ev, hipStream_t s);
hipError_t fsm_launch_src3d(const Event3D &ev, hipStream_t s);
hipError_t fsm_launch_stream3d(const Event3D &ev, hipStream_t s);
hipError_t fsm_launch_transpose(const float *in, float *out, int rows, int cols, int planes, long in_plane,
                                long out_plane, hipStream_t s);
hipError_t fsm_launch_landau(const EventLandau &ev, hipStream_t s);
// One per-cell stage (pressure, bulk pressure or shear tensor) on given planes.
hipError_t fsm_launch_stage_cells(const StageCells &a, hipStream_t s);
// Raises the dynamic-LDS ceiling of the stream kernels (bytes per workgroup).
hipError_t fsm_set_lds_limits(int bytes2d, int bytes3d);

// 2-D stream kernel variants (fsm_stream2d.hip): tile size in cells and the
// dynamic LDS a launch needs for a knot box of box_floats4 float4 elements.
int fsm_stream2d_tile_w(int variant);
int fsm_stream2d_tile_h(int variant);
size_t fsm_stream2d_lds_bytes(int variant, int box_floats4);
// Dynamic LDS of one event's stream tiles (FP32 or MFMA form, per ev.mfma).
size_t fsm_stream2d_event_lds_bytes(const Event2D &ev);
hipError_t fsm_stream2d_set_lds_limit(int bytes);



// Batched launchers: `items` are int4 work items in device memory,
// `evs` device arrays of event descriptors (see fsm_api.cpp, fsm_batch).
hipError_t fsm_launch_spline2d_batch(const Event2D *evs, const int4 *items, int n, hipStream_t s);
// One item (event, -, -) per 2-D event of the MFMA form: its knot scale.
hipError_t fsm_launch_knot_scale_batch(const Event2D *evs, const int4 *items, int n, hipStream_t s);
hipError_t fsm_launch_stream2d_batch(const Event2D *evs, const int4 *items, int n, size_t lds_bytes, int mfma,
                                     hipStream_t s);
hipError_t fsm_launch_src3d_batch(const Event3D *evs, const int4 *items, int n, hipStream_t s);
hipError_t fsm_launch_stream3d_batch(const Event3D *evs, const int4 *items, int n, size_t lds_bytes, hipStream_t s);
// Dynamic LDS of the 3-D Landau tiles (17 output planes, 22 with baryon).
size_t fsm_landau_lds_bytes(int baryon);
// 3-D items (event, transverse tile, eta tile, 1) of events with eta-major T.
hipError_t fsm_launch_landau_batch(const EventLandau *evs, const int4 *items, int n, size_t lds_bytes,
                                   hipStream_t s);
// 2-D items of events with eta_zero set (the 3x3 Landau frame).
hipError_t fsm_launch_landau2d_batch(const EventLandau *evs, const int4 *items, int n, hipStream_t s);
// Diagnostics (fsm_diag.hip).
hipError_t fsm_launch_diagnostics(const fsm_params &p, const float *e, const float *u, const float *pr,
                                  const float *Pi, const float *pi, const fsm_diag_out &out, hipStream_t s);
#endif  // FSM_KERNELS_H_
