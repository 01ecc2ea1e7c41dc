// fsm_kernels.h -- host-side launchers of the kernels in fsm_kernels.hip.
#ifndef FSM_KERNELS_H_
#define FSM_KERNELS_H_

#include <hip/hip_runtime.h>

#include "fsm_internal.h"

hipError_t fsm_launch_spline2d(const Event2D &ev, hipStream_t s);
hipError_t fsm_launch_stream2d(const Event2D &ev, hipStream_t s);
hipError_t fsm_launch_src3d(const Event3D &ev, hipStream_t s);
hipError_t fsm_launch_stream3d(const Event3D &ev, hipStream_t s);
hipError_t fsm_launch_transpose(const float *in, float *out, int rows, int cols, int planes, long in_plane,
                                long out_plane, hipStream_t s);
hipError_t fsm_launch_landau(const EventLandau &ev, hipStream_t s);
// Raises the dynamic-LDS ceiling of the stream kernels (bytes per workgroup).
hipError_t fsm_set_lds_limits(int bytes2d, int bytes3d);

#endif  // FSM_KERNELS_H_
