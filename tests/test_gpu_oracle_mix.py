"""Batched (C3) and served (C4) events against the oracle's reference float
path.  The per-event result of run_freestream_milne()
(src/FreestreamMilne.cpp:139-741, output pack :677-718) is the bar for every
event of a batch and every request of a serving mix, with the tolerances of
tests/parity.py, on both quadrature paths.

C3: a batch of 2-D events of mixed shapes (side 101-301, N_phi 48-200 before
the N_phi rule, DTAU/DX not whole numbers) in one graph-replayed launch per
stage; and the full 512-event batch on the mixed path, every fourth event
against the oracle.
C4: a serving mix of 2-D events, C2-style 3-D events (DIM_ETA from 41 to 101;
the largest on a narrower transverse window so that the CPU oracle finishes
in seconds) and baryon-carrying requests (n_B and V^mu against the oracle
too), submitted while earlier requests run."""
import time

import numpy as np
import pytest

import fsm_amd
import fsm_amd.workloads as W
import oracle as O
from parity import TOL, condition_numbers, excluded_cells_check, output_errs, well_conditioned

pytestmark = pytest.mark.gpu

_REFS = {}


def _ref(key, q, e, nb=None, fp64=True):
    """The oracle's float path, its condition numbers and (fp64) the FP64
    oracle's outputs for one event (cached across the two paths)."""
    if key not in _REFS:
        r = O.run(q, e, "float", nB=nb)
        d = O.run(q, e, "double", want_extra=False) if fp64 else None
        _REFS[key] = (r, condition_numbers(r["T"], q.TAU), d)
    return _REFS[key]


def _check(key, q, e, out, path, nb=None, nb_out=None, V_out=None, fp64=True):
    """The strict bar on the well-conditioned cells; on the others (with
    fp64) the GPU within twice the float path's own error against FP64."""
    r, kappa, d = _ref(key, q, e, nb, fp64)
    cond = well_conditioned(r["T"], q.TAU, path, kappa)
    errs = output_errs(out, r["out"], np.abs(r["T"][0]).max(), cond)
    for k in ("e", "p", "u", "pi", "Pi"):
        assert errs[k] <= TOL, (key, k, errs[k])
    if d is not None:
        errs["left_out"] = excluded_cells_check(out, r["out"], d["out"], cond, q, path, r["T"], d["T"])
    if nb is not None:
        e_ref = r["out"][0]
        m = (e_ref > 1e-3 * e_ref.max()) & cond
        scale = np.abs(r["nB"]).max()
        errs["nB"] = np.abs(nb_out - r["nB"])[m].max() / scale
        errs["V"] = max(np.abs(V_out[c] - r["V"][c])[m].max() for c in range(4)) / scale
        assert errs["nB"] <= TOL and errs["V"] <= TOL, (key, errs)
    return errs


@pytest.fixture(scope="module", params=["fp32", "mixed"])
def ctx(request):
    c = fsm_amd.Context(0)
    c.set_path(fsm_amd.PATH_MIXED if request.param == "mixed" else fsm_amd.PATH_FP32)
    c.path_name = request.param
    yield c
    c.close()


def _c3_events(n, seed):
    evs = [(fsm_amd.params_finalize(p, e), e) for p, e in W.batch_mix(n, seed=seed)]
    return evs


def test_c3_mix_contains_the_hard_shapes():
    evs = _c3_events(32, 5)
    sides = {q.DIM_X for q, _ in evs}
    assert 301 in sides and 101 in sides
    assert any(q.DIM_PHIP == 200 for q, _ in evs)
    assert any(q.DIM_X == 301 and q.DIM_PHIP >= 100 for q, _ in evs)
    # DTAU/DX is not a whole number for any event (generic shifts)
    assert all(abs(q.DTAU / q.DX - round(q.DTAU / q.DX)) > 1e-3 for q, _ in evs)


def _run_batch(ctx, evs):
    import torch

    s = torch.cuda.current_stream()
    ins = [torch.from_numpy(e).cuda() for _, e in evs]
    outs = [torch.full((17 * q.DIM,), -1.0, dtype=torch.float64, device="cuda") for q, _ in evs]
    b = fsm_amd.Batch(ctx)
    for (q, _), d, o in zip(evs, ins, outs):
        b.add(q, d.data_ptr(), o.data_ptr())
    for _ in range(2):  # the second run replays the captured graph
        b.run(s)
    torch.cuda.synchronize()
    b.close()
    return [o.cpu().numpy().reshape(17, -1) for o in outs]


def test_c3_batch_against_oracle(ctx):
    evs = _c3_events(32, 5)
    print("C3 batch of", len(evs), "events,", ctx.path_name, flush=True)
    outs = _run_batch(ctx, evs)
    worst = {}
    for k, ((q, e), out) in enumerate(zip(evs, outs)):
        errs = _check(("c3", k), q, e, out, ctx.path_name)
        for n in ("e", "u", "pi"):
            worst[n] = max(worst.get(n, 0.0), errs[n])
        worst["left_out_frac"] = max(worst.get("left_out_frac", 0.0), errs["left_out"]["frac"])
    print("C3", ctx.path_name, "worst over the batch:", worst, flush=True)


def test_c3_full_512_batch_against_oracle():
    """The whole 512-event mix of the benchmark in one batch (mixed path);
    every fourth event against the oracle, every event finite."""
    ctx = fsm_amd.Context(0)
    ctx.set_path(fsm_amd.PATH_MIXED)
    try:
        evs = _c3_events(512, 7)
        print("C3 512-event batch", flush=True)
        outs = _run_batch(ctx, evs)
        for out in outs:
            assert np.all(np.isfinite(out))
        worst = {}
        for k in range(0, len(evs), 4):
            q, e = evs[k]
            errs = _check(("c3full", k), q, e, outs[k], "mixed", fp64=False)
            for n in ("e", "u", "pi"):
                worst[n] = max(worst.get(n, 0.0), errs[n])
        print("C3 512 (every 4th event) worst:", worst, flush=True)
    finally:
        ctx.close()


def _c4_pool():
    """2-D C3-style events, C2-style 3-D events with DIM_ETA 41..101, and a
    quarter of the requests with n_B (BARYON = 1, SIGMA_B = 0.8)."""
    pool = []
    for k, (p, e) in enumerate(W.batch_mix(12, seed=41)):
        pool.append((p, e))
    pool.insert(2, W.c2_like(41, seed=101))
    pool.insert(6, W.c2_like(73, seed=102, nx=81))
    pool.insert(10, W.c2_like(101, seed=103, nx=61))
    out = []
    for k, (p, e) in enumerate(pool):
        nb = None
        if k in (1, 2, 7, 12):  # pool[2] is the 131^2 x 41 event
            p = W.with_baryon(p)
            nb = W.baryon_density(e)
        out.append((fsm_amd.params_finalize(p, e), e, nb))
    return out


def test_c4_served_mix_against_oracle(ctx):
    pool = _c4_pool()
    assert sum(q.DIM_ETA > 1 for q, _, _ in pool) == 3 and sum(nb is not None for _, _, nb in pool) == 4
    assert any(q.DIM_ETA > 1 and nb is not None for q, _, nb in pool)
    srv = fsm_amd.Server(ctx, max_batch_events=4, max_inflight=2)
    print("C4 served mix of", len(pool), "requests,", ctx.path_name, flush=True)
    ids = {}
    pending = list(range(len(pool)))
    # the first requests start the device; the rest arrive one per
    # completion, while the other slot's launch is still running
    for k in pending[:5]:
        q, e, nb = pool[k]
        ids[srv.submit(q, e, nb)] = k
    nxt = 5
    results = {}
    t0 = time.time()
    while len(results) < len(pool):
        rid = srv.poll(2000000)
        assert rid is not None, "no completion within 2 s"
        k = ids[rid]
        q, e, nb = pool[k]
        if nb is None:
            results[k] = (srv.fetch(rid), None, None)
        else:
            results[k] = srv.fetch_baryon(rid)
        if nxt < len(pool):
            q2, e2, nb2 = pool[nxt]
            ids[srv.submit(q2, e2, nb2)] = nxt
            nxt += 1
        assert time.time() - t0 < 90
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(pool) and st["bytes_in_use"] == 0
    for k, (q, e, nb) in enumerate(pool):
        outs, nbg, Vg = results[k]
        out = np.stack([outs[n] for n in fsm_amd.OUTPUT_NAMES])
        errs = _check(("c4", k), q, e, out, ctx.path_name, nb, nbg, Vg)
        print("C4", ctx.path_name, k, "dim", (q.DIM_X, q.DIM_ETA), "baryon" if nb is not None else "",
              {n: round(float(errs[n]), 9) for n in errs if n in ("e", "u", "pi", "nB", "V")}, errs["left_out"],
              flush=True)
