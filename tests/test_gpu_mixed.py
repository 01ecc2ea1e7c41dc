"""Mixed-precision path (FSM_PATH_MIXED): the 2-D quadrature on FP16 MFMA with
hi+lo split knots and weights (in 3-D the quadrature stays on the FP32 stream
kernel); the Landau frame is FP64 on both paths.  Same bar as the FP32 path: T^{mu nu} within 1e-4 of
max T^{tau tau} against the reference float path (oracle) and, in 2-D,
against the FP64 oracle on cells whose samples all stay inside the grid
(float and FP64 keep different samples at the mask boundary); e, u, pi, Pi
as in tests/parity.py.
Inputs with spline undershoot (single-cell spikes) exercise the tiles that
fall back to the FP32 sample loop inside the MFMA kernel."""
import numpy as np
import pytest

import fsm_amd
import fsm_amd.workloads as W
import oracle as O
from parity import (TOL, excluded_cells_check, interior_mask, output_errs, ring_interior_mask, tmunu_err, tmunu_local_err,
                    well_conditioned)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mixed_ctx():
    ctx = fsm_amd.Context(0)
    ctx.set_path(fsm_amd.PATH_MIXED)
    yield ctx
    ctx.close()


def _case(name):
    if name == "rect2d":
        p = W.base_params(DIM_X=77, DIM_Y=64, DIM_RAP=2, DIM_PHIP=40, DX=0.08, DY=0.11, TAU0=0.2, DTAU=0.7)
        return p, W.hotspots_2d(77, 64, 0.08, 0.11, 30, 11).ravel()
    if name == "spiky":  # white noise: the spline undershoots almost everywhere
        p = W.base_params(DIM_X=61, DIM_Y=67, DIM_PHIP=36, DX=0.1, DY=0.1, TAU0=0.1, DTAU=0.6)
        rng = np.random.Generator(np.random.PCG64(5))
        return p, rng.exponential(1.0, size=61 * 67).astype(np.float32)
    if name == "hotspots_spikes":  # smooth event with three spikes: some tiles fall back
        p = W.base_params(DIM_X=131, DIM_Y=131, DIM_PHIP=64, DX=0.08, DY=0.08, TAU0=0.1, DTAU=0.9)
        e = W.hotspots_2d(131, 131, 0.08, 0.08, 60, 21)
        for ix, iy in ((20, 30), (65, 65), (110, 40)):
            e[ix, iy] += 40.0
        return p, e.ravel()
    if name == "small3d":
        return W.small_3d()
    if name == "small3d_int":  # DTAU/DX integer: samples on grid lines
        return W.small_3d(nx=33, neta=17, nrap=7, nphi=16, dtau=0.5, seed=9)
    if name == "rect3d":  # DX != DY, odd N_phi, a ring wider than the grid
        p = W.base_params(DIM_X=57, DIM_Y=43, DIM_ETA=15, DIM_RAP=7, DIM_PHIP=151, DX=0.1, DY=0.08, DETA=0.12,
                          TAU0=0.1, DTAU=1.9, SIGMA=0.8)
        e2 = W.hotspots_2d(57, 43, 0.1, 0.08, 15, 31, width=0.4)
        env = W.eta_envelope(15, 0.12, flat=0.4, tail=0.4)
        return p, (e2[:, :, None] * env[None, None, :]).astype(np.float32).ravel()
    if name == "wide3d":  # more rows than one source window holds
        p = W.base_params(DIM_X=151, DIM_Y=101, DIM_ETA=7, DIM_RAP=5, DIM_PHIP=20, DX=0.1, DY=0.1, DETA=0.1,
                          TAU0=0.1, DTAU=2.0, SIGMA=1.0)
        e2 = W.hotspots_2d(151, 101, 0.1, 0.1, 40, 17, width=0.6)
        return p, (e2[:, :, None] * np.ones((1, 1, 7))).astype(np.float32).ravel()
    if name == "trento_ref":  # the reference's own example TRENTO event (tests/golden/trento_e.dat)
        from test_gpu_parity import reference_trento_event

        return reference_trento_event()
    if name in ("tiny2d", "small2d"):  # grids smaller than one spline or stream tile
        from test_gpu_parity import _case as parity_case

        return parity_case(name)
    if name == "int2d":  # DTAU/DX = 10: samples on grid lines next to the mask edge
        p = W.base_params(DIM_X=101, DIM_Y=101, DIM_PHIP=64, DX=0.08, DY=0.08, TAU0=0.1, DTAU=0.8)
        return p, W.hotspots_2d(101, 101, 0.08, 0.08, 60, 13).ravel()
    if name == "lastline2d":  # kept samples on the last knot line take the reference's bilinear
        from test_gpu_parity import last_knot_line_event

        return last_knot_line_event()
    return W.config(name)


def _gpu(ctx, q, e):
    import torch

    D = q.DIM
    d_e = torch.from_numpy(np.ascontiguousarray(e, dtype=np.float32)).cuda()
    d_T = torch.empty(10 * D, dtype=torch.float32, device="cuda")
    d_out = torch.empty(17 * D, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream()
    ctx.stream_tmunu_device(q, d_e.data_ptr(), d_T.data_ptr(), s)
    ctx.run_device(q, d_e.data_ptr(), d_out.data_ptr(), s)
    torch.cuda.synchronize()
    return d_T.cpu().numpy().reshape(10, D), d_out.cpu().numpy().reshape(17, D)


@pytest.mark.parametrize("name", ["C0", "C1", "rect2d", "int2d", "spiky", "hotspots_spikes", "trento_ref", "small3d",
                                  "small3d_int", "rect3d", "wide3d", "lastline2d", "tiny2d", "small2d"])
def test_mixed_parity(mixed_ctx, name):
    p, e = _case(name)
    q = fsm_amd.params_finalize(p, e)
    print("mixed case", name, "DIM", q.DIM, "DIM_PHIP", q.DIM_PHIP, flush=True)
    T, out = _gpu(mixed_ctx, q, e)
    ref = O.run(q, e, "float")
    dref = O.run(q, e, "double")
    tf = tmunu_err(T, ref["T"])
    tl = tmunu_local_err(T, ref["T"], ref["out"])
    print(name, "mixed T err per cell (relative to the cell's T^tt) %.3g" % tl, flush=True)
    assert tl <= TOL, tl
    inner = ring_interior_mask(q) if q.DIM_ETA == 1 else interior_mask(q)
    td = tmunu_err(T, dref["T"], keep=inner)
    T00max = np.abs(ref["T"][0]).max()
    cond = well_conditioned(ref["T"], q.TAU, "mixed")
    errs = output_errs(out, ref["out"], T00max, cond)
    errs64 = output_errs(out, dref["out"], np.abs(dref["T"][0]).max(), cond & inner)
    print(name, "mixed T err vs float path %.3g, vs FP64 (cells whose samples stay inside) %.3g" % (tf, td),
          flush=True)
    print("  vs float path", errs, flush=True)
    print("  vs FP64      ", errs64, flush=True)
    # rect3d streams over a ring wider than the grid: most cells see one side
    exc = excluded_cells_check(out, ref["out"], dref["out"], cond, q, "mixed", ref["T"], dref["T"],
                               budget=0.75 if name == "rect3d" else None)
    print("  cells left out of the e/u bar, against FP64:", exc, flush=True)
    assert tf <= TOL
    if q.DIM_ETA == 1:
        # In 3-D the float path itself differs from FP64 where a sample sits on
        # an eta grid line (the central rapidity node); T there is the FP32
        # path's T (tests/test_oracle_pipeline.py), only the Landau stage differs.
        assert td <= TOL
    for k in ("e", "p", "u", "pi", "Pi"):
        assert errs[k] <= TOL, (k, errs[k])
        if q.DIM_ETA == 1:
            assert errs64[k] <= TOL, (k, errs64[k])


def test_mixed_parity_c2_full(mixed_ctx):
    """The 131^2 x 101 3-D event (DIM_RAP 21, N_phi 51) on the mixed path."""
    p, e = W.config("C2")
    q = fsm_amd.params_finalize(p, e)
    print("mixed case C2 DIM", q.DIM, "DIM_PHIP", q.DIM_PHIP, flush=True)
    T, out = _gpu(mixed_ctx, q, e)
    ref = O.run(q, e, "float")
    tf = tmunu_err(T, ref["T"])
    errs = output_errs(out, ref["out"], np.abs(ref["T"][0]).max(), well_conditioned(ref["T"], q.TAU, "mixed"))
    print("C2 mixed T err vs float path %.3g" % tf, errs, flush=True)
    assert tf <= TOL
    for k in ("e", "p", "u", "pi", "Pi"):
        assert errs[k] <= TOL, (k, errs[k])


def test_mixed_runs_the_mfma_form(mixed_ctx, gpu_ctx):
    """C1 has no spline undershoot, so every tile takes the MFMA form: the
    result differs from the FP32 path in rounding, and only there."""
    p, e = W.config("C1")
    q = fsm_amd.params_finalize(p, e)
    Tm, _ = _gpu(mixed_ctx, q, e)
    Tf, _ = _gpu(gpu_ctx, q, e)
    d = np.abs(Tm.astype(np.float64) - Tf)
    frac = (d[0] > 0).mean()
    print("C1 mixed vs FP32: max rel %.3g, cells differing %.3f" % (d.max() / np.abs(Tf[0]).max(), frac))
    assert frac > 0.5
    assert d.max() <= TOL * np.abs(Tf[0]).max()


def test_mixed_batch_matches_single_events(mixed_ctx):
    import torch

    evs = [(fsm_amd.params_finalize(p, e), e) for p, e in W.batch_mix(5, seed=4)]
    q, e = _case("hotspots_spikes")
    evs.append((fsm_amd.params_finalize(q, e), e))
    q, e = W.small_3d()
    evs.append((fsm_amd.params_finalize(q, e), e))
    mixed_ctx.set_tile_policy(fsm_amd.TILES_BATCH)  # one-event calls use the batch tile shape
    for q, _ in evs:
        mixed_ctx.autotune(q)
    s = torch.cuda.current_stream()
    ins = [torch.from_numpy(np.ascontiguousarray(e, dtype=np.float32)).cuda() for _, e in evs]
    singles = []
    for (q, _), d in zip(evs, ins):
        o = torch.empty(17 * q.DIM, dtype=torch.float64, device="cuda")
        mixed_ctx.run_device(q, d.data_ptr(), o.data_ptr(), s)
        singles.append(o)
    torch.cuda.synchronize()
    singles = [o.cpu().numpy() for o in singles]
    b = fsm_amd.Batch(mixed_ctx)
    outs = []
    for (q, _), d in zip(evs, ins):
        o = torch.full((17 * q.DIM,), -1.0, dtype=torch.float64, device="cuda")
        b.add(q, d.data_ptr(), o.data_ptr())
        outs.append(o)
    for replay in range(3):
        for o in outs:
            o.fill_(-7.0)
        torch.cuda.synchronize()
        b.run(s)
        torch.cuda.synchronize()
        for k, (o, ref) in enumerate(zip(outs, singles)):
            assert np.array_equal(o.cpu().numpy(), ref), (replay, k)
    b.close()
    mixed_ctx.set_tile_policy(fsm_amd.TILES_AUTO)


@pytest.mark.parametrize("name", ["C1", "hotspots_spikes"])
def test_mixed_landau_on_identical_input(mixed_ctx, name):
    """Stage B alone on the mixed path against the oracle's FP64 eigen-solve on the
    oracle's float T^{mu nu}, every cell with e > 1e-3 e_max."""
    import torch

    p, e = _case(name)
    q = fsm_amd.params_finalize(p, e)
    ref = O.run(q, e, "float")
    D = q.DIM
    T32 = ref["T"].astype(np.float32)
    d_T = torch.from_numpy(np.ascontiguousarray(T32.ravel())).cuda()
    d_e = torch.empty(D, dtype=torch.float32, device="cuda")
    d_u = torch.empty(4 * D, dtype=torch.float32, device="cuda")
    mixed_ctx.landau_device(q, d_T.data_ptr(), d_e.data_ptr(), d_u.data_ptr(), stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    eg = d_e.cpu().numpy().astype(np.float64)
    ug = d_u.cpu().numpy().reshape(4, D).astype(np.float64)
    eo, uo = O.landau(T32.astype(np.float64).T, np.float32(q.TAU))
    m = eo > 1e-3 * eo.max()
    ee = (np.abs(eg - eo)[m] / eo[m]).max()
    ue = (np.abs(ug.T - uo)[m].max(axis=1) / uo[m, 0]).max()
    print(name, "mixed stage B: e rel err %.3g, u rel err %.3g" % (ee, ue), flush=True)
    assert ee < 1e-5 and ue < 1e-5
