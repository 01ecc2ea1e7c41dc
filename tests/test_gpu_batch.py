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
    """With the batch tile policy a one-event call uses the same tile shape as
    the batch (the autotuner's choice), so the results are bitwise equal."""
    evs = _events()
    gpu_ctx.set_tile_policy(fsm_amd.TILES_BATCH)
    try:
        _batch_vs_single(gpu_ctx, evs)
    finally:
        gpu_ctx.set_tile_policy(fsm_amd.TILES_AUTO)


def _small_2d_events():
    """2-D grids smaller than one spline tile (16 x 32 knots) or one stream
    tile, with rings of 1.7 to 4 cells: the batch spline kernel's bands and
    windows reach past every edge of the grid.  Hotspots on a floor of a
    tenth of the peak (no nearly null tensors in the edge cells)."""
    evs = []
    for nx, ny, dx, dtau, nphi, seed in ((21, 23, 0.1, 0.4, 24, 21), (9, 13, 0.3, 0.5, 16, 22), (17, 40, 0.12, 0.3, 20, 23)):
        p = W.base_params(DIM_X=nx, DIM_Y=ny, DIM_PHIP=nphi, DX=dx, DY=dx, TAU0=0.3, DTAU=dtau)
        e = W.hotspots_2d(nx, ny, dx, dx, 12, seed, width=0.4).ravel()
        e = e + np.float32(0.1) * e.max()
        evs.append((fsm_amd.params_finalize(p, e), e))
    return evs


@pytest.mark.parametrize("path", ["fp32", "mixed"])
def test_batch_small_grids_match_single_events(gpu_ctx, path):
    evs = _small_2d_events()
    gpu_ctx.set_path(fsm_amd.PATH_MIXED if path == "mixed" else fsm_amd.PATH_FP32)
    gpu_ctx.set_tile_policy(fsm_amd.TILES_BATCH)
    try:
        _batch_vs_single(gpu_ctx, evs)
    finally:
        # the session's shared context goes back to its defaults
        gpu_ctx.set_tile_policy(fsm_amd.TILES_AUTO)
        gpu_ctx.set_path(fsm_amd.PATH_FP32)


def _batch_vs_single(gpu_ctx, evs):
    import torch

    for q, _ in evs:
        gpu_ctx.autotune(q)
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


def test_batch_within_tolerance_of_latency_shape(gpu_ctx):
    """Default policy: one-event calls use the latency tile shape, batches the
    tuned shape; only the summation order differs (e within 1e-5 relative on
    cells with e > 1e-3 e_max)."""
    import torch

    evs = _events()
    s = torch.cuda.current_stream()
    b = fsm_amd.Batch(gpu_ctx)
    ins = [torch.from_numpy(e).cuda() for _, e in evs]
    outs, singles = [], []
    for (q, _), d in zip(evs, ins):
        o = torch.empty(17 * q.DIM, dtype=torch.float64, device="cuda")
        b.add(q, d.data_ptr(), o.data_ptr())
        outs.append(o)
        o1 = torch.empty(17 * q.DIM, dtype=torch.float64, device="cuda")
        gpu_ctx.run_device(q, d.data_ptr(), o1.data_ptr(), s)
        singles.append(o1)
    b.run(s)
    torch.cuda.synchronize()
    for k, (o, o1) in enumerate(zip(outs, singles)):
        a, r = o.cpu().numpy().reshape(17, -1), o1.cpu().numpy().reshape(17, -1)
        m = r[0] > 1e-3 * r[0].max()
        err = (np.abs(a[0] - r[0])[m] / r[0][m]).max()
        assert err < 1e-5, (k, err)
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


def test_fp32_batch_of_large_3d_events_matches_single_events():
    """More 3-D Landau tiles than the batch launch has workgroups (four
    131^2 x 101 events: 4 x 7511 tiles against 16384 workgroups), so each
    workgroup runs several tiles in a row, on the FP32 path, whose Landau solve
    has no barrier of its own.  Every event must equal the same event run
    alone, bitwise."""
    import torch

    ctx = fsm_amd.Context(0)
    ctx.set_path(fsm_amd.PATH_FP32)
    try:
        evs = []
        for seed in (2, 3, 4, 5):
            p, e = W.c2_like(101, seed=seed)
            evs.append((fsm_amd.params_finalize(p, e), e))
        print("FP32 batch of four 131^2x101 events", flush=True)
        s = torch.cuda.current_stream()
        ins = [torch.from_numpy(e).cuda() for _, e in evs]
        singles = []
        for (q, _), d in zip(evs, ins):
            o = torch.empty(17 * q.DIM, dtype=torch.float64, device="cuda")
            ctx.run_device(q, d.data_ptr(), o.data_ptr(), s)
            torch.cuda.synchronize()
            singles.append(o.cpu())
            del o
        b = fsm_amd.Batch(ctx)
        outs = [torch.full((17 * q.DIM,), -1.0, dtype=torch.float64, device="cuda") for q, _ in evs]
        for (q, _), d, o in zip(evs, ins, outs):
            b.add(q, d.data_ptr(), o.data_ptr())
        for replay in range(2):
            b.run(s)
            torch.cuda.synchronize()
            for k, (o, ref) in enumerate(zip(outs, singles)):
                assert torch.equal(o.cpu(), ref), (replay, k)
        b.close()
    finally:
        ctx.close()


def test_calls_on_two_streams_are_ordered(gpu_ctx):
    """Two same-shape events on two streams of one context share the
    context's workspace; the second call waits for the first, so each result
    equals the event run alone."""
    import torch

    pa, ea = W.config("C1")
    pb, eb = W.config("C1", seed=9)
    qa, qb = fsm_amd.params_finalize(pa, ea), fsm_amd.params_finalize(pb, eb)
    s0 = torch.cuda.current_stream()
    da, db = torch.from_numpy(ea).cuda(), torch.from_numpy(eb).cuda()
    refs = []
    for q, d in ((qa, da), (qb, db)):
        o = torch.empty(17 * q.DIM, dtype=torch.float64, device="cuda")
        gpu_ctx.run_device(q, d.data_ptr(), o.data_ptr(), s0)
        torch.cuda.synchronize()
        refs.append(o.cpu())
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for rep in range(3):
        oa = torch.empty(17 * qa.DIM, dtype=torch.float64, device="cuda")
        ob = torch.empty(17 * qb.DIM, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        for _ in range(4):
            gpu_ctx.run_device(qa, da.data_ptr(), oa.data_ptr(), s1)
            gpu_ctx.run_device(qb, db.data_ptr(), ob.data_ptr(), s2)
        torch.cuda.synchronize()
        assert torch.equal(oa.cpu(), refs[0]) and torch.equal(ob.cpu(), refs[1]), rep


@pytest.mark.parametrize("path", ["fp32", "mixed"])
def test_cost_estimate_tracks_measured_batch_time(path):
    """The serving loop caps each launch by the plan's estimate of an event's
    device time inside a batch (fsm_plan_info.est_batch_us); it must be the
    measured time to within a factor of three, for 2-D and 3-D events."""
    import time

    import torch

    ctx = fsm_amd.Context(0)
    ctx.set_path(fsm_amd.PATH_MIXED if path == "mixed" else fsm_amd.PATH_FP32)
    try:
        for name, B in (("C1", 64), ("C2", 2)):
            p, e = W.config(name)
            q = fsm_amd.params_finalize(p, e)
            d_e = torch.from_numpy(e).cuda()
            outs = [torch.empty(17 * q.DIM, dtype=torch.float64, device="cuda") for _ in range(B)]
            b = fsm_amd.Batch(ctx)
            for o in outs:
                b.add(q, d_e.data_ptr(), o.data_ptr())
            s = torch.cuda.current_stream()
            for _ in range(3):
                b.run(s)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = 10
            for _ in range(reps):
                b.run(s)
            torch.cuda.synchronize()
            measured = (time.perf_counter() - t0) / (reps * B) * 1e6
            est = ctx.plan_info(q)["est_batch_us"]
            print(name, path, "estimated %.1f us per event, measured %.1f us" % (est, measured), flush=True)
            assert measured / 3 <= est <= 3 * measured, (name, est, measured)
            b.close()
    finally:
        ctx.close()
