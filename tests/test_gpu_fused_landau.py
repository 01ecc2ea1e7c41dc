"""The 2-D Landau stage fused into the stream kernels' tile epilogue (events
without a baryon current: one-event runs, batches, the server) against the
stage as a separate launch on the same tensor: fsm_stream_tmunu_device
writes T^{mu nu}, fsm_landau_device solves it (src/LandauMatch.cpp:145-367,
src/EquationOfState.cpp:5), and every output of fsm_run_device must be the
same bits, on both paths -- including the MFMA kernel's FP32 tiles where the
spline undershoots ("spiky", "hotspots_spikes") and an event whose samples
sit on grid lines ("int2d").  fsm_landau_device takes the caller's tensor
through the 4x4 frame (the quartic) and the fused stage through the 3x3 one
(the cubic); on these tensors they agree to the bit except for the sign of a
zero u^eta (the quartic's eigenvector can give -0), which is compared by
value."""
import numpy as np
import pytest

import fsm_amd
import fsm_amd.workloads as W

pytestmark = pytest.mark.gpu

HBARC = np.float64(np.float32(0.197326938))


def _case(name):
    if name == "C1":
        return W.config("C1")
    from test_gpu_parity import _case as parity_case

    return parity_case(name)


@pytest.mark.parametrize("path", ["fp32", "mixed"])
@pytest.mark.parametrize("name", ["C1", "spiky", "hotspots_spikes", "int2d", "rect2d"])
def test_fused_landau_matches_separate_launch(name, path):
    import torch

    p, e = _case(name)
    ctx = fsm_amd.Context(0)
    ctx.set_path(fsm_amd.PATH_MIXED if path == "mixed" else fsm_amd.PATH_FP32)
    q = fsm_amd.params_finalize(p, e)
    assert q.DIM_ETA == 1
    D = q.DIM
    s = torch.cuda.current_stream()
    d_e = torch.from_numpy(np.ascontiguousarray(e, dtype=np.float32).ravel()).cuda()
    d_out = torch.full((17 * D,), float("nan"), dtype=torch.float64, device="cuda")
    ctx.run_device(q, d_e.data_ptr(), d_out.data_ptr(), s)  # fused
    d_T = torch.empty(10 * D, dtype=torch.float32, device="cuda")
    f = {k: torch.empty(n * D, dtype=torch.float32, device="cuda") for k, n in
         (("e", 1), ("u", 4), ("p", 1), ("Pi", 1), ("pi", 10))}
    ctx.stream_tmunu_device(q, d_e.data_ptr(), d_T.data_ptr(), s)
    ctx.landau_device(q, d_T.data_ptr(), f["e"].data_ptr(), f["u"].data_ptr(), f["p"].data_ptr(),
                      f["Pi"].data_ptr(), f["pi"].data_ptr(), stream=s)
    torch.cuda.synchronize()
    ctx.close()
    out = d_out.cpu().numpy().reshape(17, D)
    assert not np.isnan(out).any()
    sep = {k: v.cpu().numpy().astype(np.float64).reshape(-1, D) for k, v in f.items()}
    # output_to_vectors order: e, p (x hbarc), u^mu, pi^{mu nu} (x hbarc), Pi (x hbarc);
    # (double)x * hbarc is exact, so the float of the fused run is recovered exactly
    want = np.concatenate([sep["e"] * HBARC, sep["p"] * HBARC, sep["u"], sep["pi"] * HBARC, sep["Pi"] * HBARC])
    # bitwise, signed zeros included, except u^eta (plane 5: by value)
    same = [np.array_equal(out[c], want[c]) if c == 5 else
            np.array_equal(out[c].view(np.int64), want[c].view(np.int64)) for c in range(17)]
    assert not out[5].any()  # 2-D: u^eta is zero
    n_diff = int(sum((out[c].view(np.int64) != want[c].view(np.int64)).sum() for c in range(17) if c != 5))
    print(name, path, "fused vs separate Landau launch: planes identical %d / 17, differing values %d"
          % (sum(same), n_diff), flush=True)
    for c in range(17):
        if not same[c]:
            d = out[c].view(np.int64) != want[c].view(np.int64)
            i = np.nonzero(d)[0][:4]
            print("  plane", c, "cells", int(d.sum()), "fused", out[c][i], "separate", want[c][i], flush=True)
    assert all(same)
