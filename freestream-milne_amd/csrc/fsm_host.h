// fsm_host.h -- internal host-side entry points shared between fsm_api.cpp
// (contexts, plans, batches) and fsm_server.cpp (the request loop).
#ifndef FSM_HOST_H_
#define FSM_HOST_H_

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "../../include/fsm.h"

namespace fsm_host {

// Device workspace one event needs inside a batch (spline or source planes
// plus the stress tensor; twice that for a baryon event's two passes).
size_t workspace_bytes(const fsm_params &p, bool baryon);
// fsm_batch_add (d_nb == NULL) or fsm_batch_add_baryon with caller-owned
// workspace of workspace_bytes(*p, d_nb != NULL).
int batch_add_ws(fsm_batch *b, const fsm_params *p, const float *d_e, const float *d_nb, double *d_out, void *ws);
// Descriptor uploads through pinned staging, ordered on the launch stream.
void batch_set_async_upload(fsm_batch *b, bool on);
// Builds (if changed) and launches the batch on `s` without a graph.
int batch_launch_now(fsm_batch *b, hipStream_t s);
int ctx_device(const fsm_ctx *ctx);
// Builds/looks up the plan for p (returns its status: unsupported shapes fail
// here) and estimates the event's device time inside a batch on the
// context's path, in microseconds (cost_us may be NULL).
int plan_check(fsm_ctx *ctx, const fsm_params *p, bool baryon, float *cost_us);

}  // namespace fsm_host

#endif  // FSM_HOST_H_
