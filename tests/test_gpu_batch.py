"""Batched execution (fsm_batch_*): a mix of event shapes, 2-D and 3-D, in
one launch per stage, replayed from a HIP graph.  Each event's result must be
identical to the same event run alone (same kernels, same arithmetic), on
every graph replay, with other GPU work and synchronisations between replays."""
import numpy as np
import pytest

import fsm_amd
import fsm_amd.workloads as W

pytestmark = pytest.mark.gpu


def _events():
    evs = W.batch_mix(6, seed=3)
    evs.append(W.config("C0"))
    evs.append(W.small_3d())
    evs.append(W.small_3d(nx=33, neta=17, nrap=7, nphi=16, dtau=0.5, seed=9))
    return [(fsm_amd.params_finalize(p, e), e) for p, e in evs]


def test_batch_matches_single_events_over_graph_replays(gpu_ctx):
    import torch

    evs = _events()
    s = torch.cuda.current_stream()
    ins = [torch.from_numpy(e).cuda() for _, e in evs]
    singles = []
    for (q, _), d in zip(evs, ins):
        o = torch.empty(17 * q.DIM, dtype=torch.float64, device="cuda")
        gpu_ctx.run_device(q, d.data_ptr(), o.data_ptr(), s)
        singles.append(o)
    torch.cuda.synchronize()
    singles = [o.cpu().numpy() for o in singles]

    b = fsm_amd.Batch(gpu_ctx)
    outs = []
    for (q, _), d in zip(evs, ins):
        o = torch.full((17 * q.DIM,), -1.0, dtype=torch.float64, device="cuda")
        assert b.add(q, d.data_ptr(), o.data_ptr()) == len(outs)
        outs.append(o)
    assert len(b) == len(evs)
    for replay in range(4):
        for o in outs:
            o.fill_(-7.0)  # eager work between replays
        torch.cuda.synchronize()
        b.run(s)
        torch.cuda.synchronize()
        for k, (o, ref) in enumerate(zip(outs, singles)):
            got = o.cpu().numpy()
            assert np.array_equal(got, ref), (replay, k)
    # eager (no graph) gives the same
    b.set_graph(False)
    b.run(s)
    torch.cuda.synchronize()
    for o, ref in zip(outs, singles):
        assert np.array_equal(o.cpu().numpy(), ref)
    b.close()


def test_batch_clear_and_reuse(gpu_ctx):
    import torch

    evs = _events()[:3]
    s = torch.cuda.current_stream()
    b = fsm_amd.Batch(gpu_ctx)
    ins = [torch.from_numpy(e).cuda() for _, e in evs]
    outs = [torch.empty(17 * q.DIM, dtype=torch.float64, device="cuda") for q, _ in evs]
    for (q, _), d, o in zip(evs, ins, outs):
        b.add(q, d.data_ptr(), o.data_ptr())
    b.run(s)
    b.clear()
    assert len(b) == 0
    q, _ = evs[1]
    o2 = torch.empty(17 * q.DIM, dtype=torch.float64, device="cuda")
    b.add(q, ins[1].data_ptr(), o2.data_ptr())
    b.run(s)
    torch.cuda.synchronize()
    assert np.array_equal(o2.cpu().numpy(), outs[1].cpu().numpy())
    b.close()
