"""fsm_run_device_noalloc: the allocation-free variant of fsm_run_device
(workspace in a fixed device array).  Each call must give bitwise the
results of fsm_run_device on the same inputs -- two calls in a row, the
second with a smaller event, on both quadrature paths, for 2-D and 3-D
events -- and the documented refusals must hold: an event larger than the
fixed workspace (FSM_ERR_TOO_LARGE) and a shape without tables in the context
(FSM_ERR_NO_PLAN)."""
import numpy as np
import pytest

import fsm_amd
import fsm_amd.workloads as W

pytestmark = pytest.mark.gpu


def _pair(ctx, q, e, s):
    import torch

    d_e = torch.from_numpy(e).cuda()
    ref = torch.full((17 * q.DIM,), -1.0, dtype=torch.float64, device="cuda")
    got = torch.full((17 * q.DIM,), -2.0, dtype=torch.float64, device="cuda")
    ctx.run_device(q, d_e.data_ptr(), ref.data_ptr(), s)  # builds the shape's tables
    st = ctx.run_device_noalloc(q, d_e.data_ptr(), got.data_ptr(), s)
    torch.cuda.synchronize()
    assert st == fsm_amd.FSM_OK, fsm_amd.lib().fsm_status_string(st)
    return ref.cpu().numpy(), got.cpu().numpy()


@pytest.mark.parametrize("path", ["fp32", "mixed"])
@pytest.mark.parametrize("kind", ["2d", "3d"])
def test_noalloc_equals_run_device(path, kind):
    import torch

    ctx = fsm_amd.Context(0)
    ctx.set_path(fsm_amd.PATH_MIXED if path == "mixed" else fsm_amd.PATH_FP32)
    try:
        if kind == "2d":
            evs = [W.config("C1"), W.config("C0")]  # the second smaller
        else:
            evs = [W.config("C2"), W.small_3d()]
        s = torch.cuda.current_stream()
        for k, (p, e) in enumerate(evs):
            q = fsm_amd.params_finalize(p, e)
            print("noalloc", kind, path, "call", k, "DIM", q.DIM, flush=True)
            ref, got = _pair(ctx, q, e, s)
            assert np.array_equal(ref, got), (kind, path, k)
    finally:
        ctx.close()


def test_noalloc_refusals():
    import torch

    ctx = fsm_amd.Context(0)
    try:
        s = torch.cuda.current_stream()
        # a shape never seen by this context: no tables, nothing built
        p = W.base_params(DIM_X=77, DIM_Y=53, DIM_PHIP=24, DX=0.1, DY=0.1, TAU0=0.1, DTAU=0.4)
        q = fsm_amd.params_finalize(p)
        d_e = torch.ones(q.DIM, dtype=torch.float32, device="cuda")
        d_o = torch.empty(17 * q.DIM, dtype=torch.float64, device="cuda")
        assert ctx.run_device_noalloc(q, d_e.data_ptr(), d_o.data_ptr(), s) == fsm_amd.FSM_ERR_NO_PLAN
        # larger than the fixed workspace (72 B per 2-D cell): refused before anything runs
        big = fsm_amd.params_finalize(W.base_params(DIM_X=1200, DIM_Y=1200, DIM_PHIP=24, DX=0.05, DY=0.05,
                                                    TAU0=0.1, DTAU=0.2))
        b_e = torch.ones(big.DIM, dtype=torch.float32, device="cuda")  # buffers of the event's real size
        b_o = torch.empty(17 * big.DIM, dtype=torch.float64, device="cuda")
        assert ctx.run_device_noalloc(big, b_e.data_ptr(), b_o.data_ptr(), s) == fsm_amd.FSM_ERR_TOO_LARGE
        torch.cuda.synchronize()
    finally:
        ctx.close()
