"""Baryon extension on the GPU (BARYON = 1): the charge pass (n_B through
convertInitialChargeDensity and a second freeStream, src/FreeStream.cpp:353-387,
src/FreestreamMilne.cpp:386), calculateBaryonCurrent and n_B, V^mu
(src/LandauMatch.cpp:110-143, :370-392), against the oracle's reference float
path, on both quadrature paths; batched and served baryon events against the
same events run alone.

Tolerance: |dn_B| and |dV^mu| <= 1e-4 max|n_B| on the cells of tests/parity.py
(e > 1e-3 e_max, well-conditioned Landau matching); the 17 outputs as in the
other parity tests."""
import numpy as np
import pytest

import fsm_amd
import fsm_amd.workloads as W
import oracle as O
from parity import TOL, output_errs, well_conditioned

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "2d":
        p = W.base_params(DIM_X=101, DIM_Y=101, DIM_PHIP=64, DX=0.08, DY=0.08, TAU0=0.1, DTAU=0.8)
        e = W.hotspots_2d(101, 101, 0.08, 0.08, 60, 13).ravel()
    else:
        p, e = W.small_3d()
    return W.with_baryon(p), e, W.baryon_density(e)


@pytest.fixture(scope="module", params=["fp32", "mixed"])
def ctx(request):
    c = fsm_amd.Context(0)
    c.set_path(fsm_amd.PATH_MIXED if request.param == "mixed" else fsm_amd.PATH_FP32)
    c.path_name = request.param
    yield c
    c.close()


@pytest.mark.parametrize("name", ["2d", "3d"])
def test_baryon_parity(ctx, name):
    p, e, nb = _case(name)
    print("baryon case", name, flush=True)
    outs, nb_g, V_g, q, tau = ctx.run_baryon(p, e, nb)
    ref = O.run(q, e, "float", nB=nb)
    g = np.stack([outs[k] for k in fsm_amd.OUTPUT_NAMES])
    cond = well_conditioned(ref["T"], q.TAU, ctx.path_name)
    errs = output_errs(g, ref["out"], np.abs(ref["T"][0]).max(), cond)
    e_ref = ref["out"][0]
    m = (e_ref > 1e-3 * e_ref.max()) & cond
    scale = np.abs(ref["nB"]).max()
    dn = np.abs(nb_g - ref["nB"])[m].max() / scale
    dv = max(np.abs(V_g[c] - ref["V"][c])[m].max() for c in range(4)) / scale
    print(name, "n_B err %.3g, V err %.3g" % (dn, dv), errs, flush=True)
    assert dn <= TOL and dv <= TOL
    for k in ("e", "u", "pi", "Pi"):
        assert errs[k] <= TOL, (k, errs[k])


def test_baryon_batch_matches_single_events(ctx):
    import torch

    evs = []
    for name in ("2d", "3d"):
        p, e, nb = _case(name)
        evs.append((fsm_amd.params_finalize(p, e), e, nb))
    p, e = W.config("C0")
    evs.append((fsm_amd.params_finalize(p, e), e, None))
    ctx.set_tile_policy(fsm_amd.TILES_BATCH)
    try:
        for q, _, _ in evs:
            ctx.autotune(q)
        s = torch.cuda.current_stream()
        dev = [(torch.from_numpy(e).cuda(), None if nb is None else torch.from_numpy(nb).cuda()) for _, e, nb in evs]
        singles, outs = [], []
        b = fsm_amd.Batch(ctx)
        for (q, _, nb), (de, dn) in zip(evs, dev):
            planes = fsm_amd.NUM_OUTPUTS if dn is None else fsm_amd.NUM_OUTPUTS_BARYON
            o1 = torch.empty(planes * q.DIM, dtype=torch.float64, device="cuda")
            o2 = torch.full((planes * q.DIM,), -3.0, dtype=torch.float64, device="cuda")
            if dn is None:
                ctx.run_device(q, de.data_ptr(), o1.data_ptr(), s)
                b.add(q, de.data_ptr(), o2.data_ptr())
            else:
                ctx.run_baryon_device(q, de.data_ptr(), dn.data_ptr(), o1.data_ptr(), s)
                b.add_baryon(q, de.data_ptr(), dn.data_ptr(), o2.data_ptr())
            singles.append(o1)
            outs.append(o2)
        for replay in range(2):
            b.run(s)
            torch.cuda.synchronize()
            for k, (o1, o2) in enumerate(zip(singles, outs)):
                assert torch.equal(o1, o2), (replay, k)
        b.close()
    finally:
        ctx.set_tile_policy(fsm_amd.TILES_AUTO)


def test_served_baryon_requests_match_runs_alone(ctx):
    import torch

    p, e, nb = _case("2d")
    p0, e0 = W.config("C0")
    ctx.set_tile_policy(fsm_amd.TILES_BATCH)
    try:
        srv = fsm_amd.Server(ctx, max_batch_events=8, max_inflight=2)
        ids = [srv.submit(p, e, nb), srv.submit(p0, e0), srv.submit(p, e, nb)]
        assert all(i is not None for i in ids)
        done = set()
        while len(done) < 3:
            rid = srv.poll(200000)
            assert rid is not None
            done.add(rid)
        res = srv.fetch_baryon(ids[0])
        res2 = srv.fetch_baryon(ids[2])
        srv.release(ids[1])
        srv.close()
        q = fsm_amd.params_finalize(p, e)
        ref_outs, ref_nb, ref_V, _, _ = ctx.run_baryon(p, e, nb)
        for r in (res, res2):
            outs, nbg, Vg = r
            for k in fsm_amd.OUTPUT_NAMES:
                assert np.array_equal(outs[k], ref_outs[k]), k
            assert np.array_equal(nbg, ref_nb)
            for c in range(4):
                assert np.array_equal(Vg[c], ref_V[c])
        assert q.DIM == e.size
    finally:
        ctx.set_tile_policy(fsm_amd.TILES_AUTO)
