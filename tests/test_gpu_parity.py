"""GPU parity: libfsm.so (HIP, through the C ABI) against the oracle's
reference float path on the same seeded inputs, with the tolerances of
tests/parity.py; the FP64 oracle is reported beside it.  Every call goes
through libfsm.so -- there is no CPU fallback in the product path."""
import os

import numpy as np
import pytest

import fsm_amd
import fsm_amd.workloads as W
import oracle as O
from parity import TOL, excluded_cells_check, interior_mask, output_errs, tmunu_err, tmunu_local_err, well_conditioned

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "small3d":
        return W.small_3d()
    if name == "small3d_int":  # DTAU/DX integer: samples on grid lines
        return W.small_3d(nx=33, neta=17, nrap=7, nphi=16, dtau=0.5, seed=9)
    if name == "rect2d":  # DIM_X != DIM_Y, DX != DY, two identical rapidity copies
        p = W.base_params(DIM_X=77, DIM_Y=64, DIM_RAP=2, DIM_PHIP=40, DX=0.08, DY=0.11, TAU0=0.2, DTAU=0.7)
        return p, W.hotspots_2d(77, 64, 0.08, 0.11, 30, 11).ravel()
    if name in ("tiny2d", "small2d"):  # grids smaller than one spline or stream tile: every cell near an edge
        from test_gpu_batch import _small_2d_events

        q, e = _small_2d_events()[1 if name == "tiny2d" else 0]
        return q, e
    if name == "int2d":  # DTAU/DX = 10: samples on grid lines next to the mask edge; hotspots at the edges
        p = W.base_params(DIM_X=101, DIM_Y=101, DIM_PHIP=64, DX=0.08, DY=0.08, TAU0=0.1, DTAU=0.8)
        return p, W.hotspots_2d(101, 101, 0.08, 0.08, 60, 13).ravel()
    if name in ("spiky", "hotspots_spikes"):
        from test_gpu_mixed import _case as mixed_case

        return mixed_case(name)
    if name == "trento_ref":
        return reference_trento_event()
    if name == "lastline2d":
        return last_knot_line_event()
    return W.config(name)


def last_knot_line_event():
    """A grid whose last row and column are a float rounding case: with
    DY = 0.09659824 the reference's (y - ymin)/DY on the last row rounds to
    179.999985, so the samples with sin(phi) ~ 0 (a shift below half an ulp
    of y, so y_new == y_max) are kept at floor 179, and because their source
    point is not strictly inside the knot range the reference takes bilinear
    instead of the spline there (src/FreeStream.cpp:198-246).  Hotspots near
    the edges make those cells matter (e > 1e-3 e_max)."""
    n = 181
    d = float(np.float32(0.09659824))
    p = W.base_params(DIM_X=n, DIM_Y=n, DIM_PHIP=200, DX=d, DY=d, TAU0=0.1, DTAU=float(np.float32(0.6522806)))
    x = W.grid_coords(n, d).astype(np.float64)
    rng = np.random.default_rng(108)
    e = np.zeros((n, n))
    for cx, cy in list(rng.uniform(-8.6, 8.6, size=(60, 2))) + [(8.6, 0.3), (-0.4, 8.6), (8.5, 8.5)]:
        amp = rng.gamma(1.0, 2.0) + 0.5
        e += amp * np.exp(-((x[:, None] - cx) ** 2 + (x[None, :] - cy) ** 2) / (2 * 0.5 ** 2))
    return p, e.astype(np.float32).ravel()


def reference_trento_event():
    """The reference's own example profile (initial_profiles/e.dat, a 131^2
    TRENTO event, committed as tests/golden/trento_e.dat) through the TRENTO
    block reader, streamed on a 0.1 fm grid for 1 fm/c.  The reader gives the
    density in fm^-4, as for every IC_ENERGY but the vector input (here 3)."""
    p = W.base_params(DIM_X=131, DIM_Y=131, DX=0.1, DY=0.1, DIM_PHIP=64, TAU0=0.1, DTAU=1.0)
    p.IC_ENERGY = 3
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "trento_e.dat")
    return p, fsm_amd.read_profile(p, fsm_amd.PROFILE_TRENTO_BLOCK, path)


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


@pytest.mark.parametrize("name", ["C0", "C1", "rect2d", "int2d", "spiky", "hotspots_spikes", "small3d", "small3d_int",
                                  "trento_ref", "lastline2d", "tiny2d", "small2d"])
def test_parity_against_reference_float_path(gpu_ctx, name):
    p, e = _case(name)
    q = fsm_amd.params_finalize(p, e)
    print("case", name, "DIM", q.DIM, "DIM_PHIP", q.DIM_PHIP, flush=True)
    T, out = _gpu(gpu_ctx, q, e)
    ref = O.run(q, e, "float")
    dref = O.run(q, e, "double")
    tf = tmunu_err(T, ref["T"])
    tl = tmunu_local_err(T, ref["T"], ref["out"])
    td = tmunu_err(T, dref["T"], keep=interior_mask(q))
    T00max = np.abs(ref["T"][0]).max()
    cond = well_conditioned(ref["T"], q.TAU)
    errs = output_errs(out, ref["out"], T00max, cond)
    print(name, "T err vs float path %.3g (per cell %.3g), vs FP64 (interior) %.3g" % (tf, tl, td), errs, flush=True)
    assert tl <= TOL, tl
    exc = excluded_cells_check(out, ref["out"], dref["out"], cond, q, "fp32", ref["T"], dref["T"])
    print(name, "ill-conditioned cells against FP64:", exc, flush=True)
    assert tf <= TOL
    for k in ("e", "p", "u", "pi", "Pi"):
        assert errs[k] <= TOL, (k, errs[k])


def test_parity_c2_full(gpu_ctx):
    p, e = W.config("C2")
    q = fsm_amd.params_finalize(p, e)
    print("case C2 DIM", q.DIM, "DIM_PHIP", q.DIM_PHIP, flush=True)
    T, out = _gpu(gpu_ctx, q, e)
    ref = O.run(q, e, "float")
    tf = tmunu_err(T, ref["T"])
    cond = well_conditioned(ref["T"], q.TAU)
    errs = output_errs(out, ref["out"], np.abs(ref["T"][0]).max(), cond)
    print("C2 T err vs float path %.3g" % tf, errs, flush=True)
    dref = O.run(q, e, "double", want_extra=False)
    print("C2 ill-conditioned cells against FP64:",
          excluded_cells_check(out, ref["out"], dref["out"], cond, q, "fp32", ref["T"], dref["T"]), flush=True)
    assert tf <= TOL
    for k in ("e", "p", "u", "pi", "Pi"):
        assert errs[k] <= TOL, (k, errs[k])


def test_host_api_equals_device_api(gpu_ctx):
    p, e = W.config("C0")
    q = fsm_amd.params_finalize(p, e)
    outs, q2, tau = gpu_ctx.run(p, e)
    assert q2.as_dict() == q.as_dict() and tau == q.TAU
    _, dev = _gpu(gpu_ctx, q, e)
    for k, name in enumerate(fsm_amd.OUTPUT_NAMES):
        assert np.array_equal(outs[name], dev[k]), name


def test_dtau_zero_known_answer_gpu(gpu_ctx):
    p = W.base_params(DIM_X=21, DIM_Y=23, DIM_PHIP=16, DX=0.1, DY=0.1, TAU0=0.5, DTAU=0.0)
    rng = np.random.default_rng(3)
    e = rng.uniform(1.0, 5.0, 21 * 23).astype(np.float32)
    outs, q, _ = gpu_ctx.run(p, e)
    hb = float(np.float32(0.197326938))
    e0 = (e.astype(np.float64) / hb + 1e-7).reshape(21, 23)[1:20, 1:22] / float(q.TAU)
    got = outs["e"].reshape(21, 23)[1:20, 1:22] / hb
    assert np.allclose(got, e0, rtol=1e-6)
    assert np.allclose(outs["pixx"].reshape(21, 23)[1:20, 1:22] / hb, e0 / 6, rtol=1e-4, atol=1e-5 * e0.max())
    assert np.all(outs["e"].reshape(21, 23)[0] == pytest.approx(1e-3 * hb))


@pytest.mark.parametrize("variant", ["outputformat1", "ideal", "eos2", "eos3", "zero_input"])
def test_switches_match_reference(gpu_ctx, variant):
    p, e = W.config("C0")
    if variant == "outputformat1":
        p.OUTPUTFORMAT = 1
    elif variant == "ideal":
        p.VISCOUS_MATCHING = 0
    elif variant == "eos2":
        p.EOS_TYPE = 2
    elif variant == "eos3":
        p.EOS_TYPE = 3
    elif variant == "zero_input":
        e = np.zeros_like(e)
    outs, q, _ = gpu_ctx.run(p, e)
    ref = O.run(q, e, "float")
    out = np.stack([outs[n] for n in fsm_amd.OUTPUT_NAMES])
    if variant == "outputformat1":
        assert np.all(out == 0.0)
        return
    T00max = np.abs(ref["T"][0]).max()
    errs = output_errs(out, ref["out"], T00max, well_conditioned(ref["T"], q.TAU))
    for k in ("e", "p", "u", "pi", "Pi"):
        assert errs[k] <= TOL, (variant, k, errs[k])
    if variant == "eos3":
        assert np.all(out[1] == 0.0)  # the placeholder EoS gives p = 0
    if variant == "ideal":
        assert np.all(out[6:] == 0.0)


def test_size_mismatch_raises(gpu_ctx):
    p, e = W.config("C0")
    with pytest.raises(fsm_amd.FsmError) as ei:
        gpu_ctx.run(p, e[:-3])
    assert ei.value.status == fsm_amd.FSM_ERR_SIZE


def test_reference_class_mirror(gpu_ctx, tmp_path):
    # FREESTREAMMILNE usage: freestream_input in the working directory,
    # initialize_from_vector, run_freestream_milne, output_to_vectors.
    lines = ["OUTPUTFORMAT 2", "BARYON 0", "IC_ENERGY 5", "IC_BARYON 1", "ETA_WIDTH 0.5", "ETA_FLAT 0.5",
             "SIGMA 1.0", "SIGMA_B 1.0", "DIM_X 101", "DIM_Y 101", "DIM_ETA 1", "DIM_RAP 1", "DIM_PHIP 51",
             "DX 0.1", "DY 0.1", "DETA 0.1", "DRAP 0.2", "DTAU 0.5", "TAU0 0.1", "EOS_TYPE 1", "E_FREEZE 1.7",
             "VISCOUS_MATCHING 1", "E_DEP_FS 0", "E_R 1.0", "TAU_R 1.0", "ALPHA 0.0"]
    f = tmp_path / "freestream_input"
    f.write_text("\n".join(lines) + "\n")
    fs = fsm_amd.FreestreamMilne(input_file=str(f), context=gpu_ctx)
    _, e = W.config("C0")
    fs.initialize_from_vector(e)
    assert fs.run_freestream_milne() == 0
    vecs = fs.output_to_vectors()
    assert len(vecs) == 17 and vecs[0].size == 101 * 101
    assert fs.tau_LandauMatch == pytest.approx(0.6)
    p, _ = W.config("C0")
    outs, _, _ = gpu_ctx.run(p, e)
    assert np.array_equal(vecs[0], outs["e"]) and np.array_equal(vecs[16], outs["Pi"])


@pytest.mark.parametrize("name", ["C1", "small3d"])
def test_landau_stage_on_identical_input(gpu_ctx, name):
    """Stage B alone: the GPU Landau kernel on the oracle's float T^{mu nu}
    against the oracle's FP64 eigen-solve with the reference selection rules,
    on every cell with e > 1e-3 e_max (no conditioning exclusion)."""
    import torch

    p, e = _case(name)
    q = fsm_amd.params_finalize(p, e)
    ref = O.run(q, e, "float")
    D = q.DIM
    T32 = ref["T"].astype(np.float32)
    d_T = torch.from_numpy(np.ascontiguousarray(T32.ravel())).cuda()
    d_e = torch.empty(D, dtype=torch.float32, device="cuda")
    d_u = torch.empty(4 * D, dtype=torch.float32, device="cuda")
    gpu_ctx.landau_device(q, d_T.data_ptr(), d_e.data_ptr(), d_u.data_ptr(), stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    eg = d_e.cpu().numpy().astype(np.float64)
    ug = d_u.cpu().numpy().reshape(4, D).astype(np.float64)
    eo, uo = O.landau(T32.astype(np.float64).T, np.float32(q.TAU))
    m = eo > 1e-3 * eo.max()
    ee = (np.abs(eg - eo)[m] / eo[m]).max()
    ue = (np.abs(ug.T - uo)[m].max(axis=1) / uo[m, 0]).max()
    fb = (eo == 1e-3) != (eg == np.float32(1e-3))
    print(name, "stage B: e rel err %.3g, u rel err %.3g, fallback mismatches %d" % (ee, ue, fb.sum()), flush=True)
    assert ee < 1e-5 and ue < 1e-5


@pytest.mark.parametrize("path", ["fp32", "mixed"])
@pytest.mark.parametrize("name", ["C1", "hotspots_spikes", "spiky", "int2d"])
def test_landau_2d_frame_on_gpu_tensor(name, path):
    """Stage B as the run path does it for 2-D events (the 3x3 frame from the
    characteristic cubic, both paths): e and u of every cell with e > 1e-3
    e_max against the oracle's FP64 eigen-solve of the GPU's own T^{mu nu}
    (no conditioning exclusion), and the reference's fallback decision on
    every cell."""
    p, e = _case(name)
    q = fsm_amd.params_finalize(p, e)
    ctx = fsm_amd.Context(0)
    ctx.set_path(fsm_amd.PATH_MIXED if path == "mixed" else fsm_amd.PATH_FP32)
    T, out = _gpu(ctx, q, e)
    ctx.close()
    hb = np.float64(np.float32(0.197326938))
    eg = out[0] / hb  # (double)e * hbarc is exact, so this is the float e
    ug = out[2:6]
    eo, uo = O.landau(T.astype(np.float64).T, np.float32(q.TAU))
    m = eo > 1e-3 * eo.max()
    ee = (np.abs(eg - eo)[m] / eo[m]).max()
    ue = (np.abs(ug.T - uo)[m].max(axis=1) / uo[m, 0]).max()
    fo = (eo == 1e-3) & (uo[:, 0] == 1.0)
    fg = (eg == np.float32(1e-3)) & (ug[0] == 1.0)
    print(name, path, "2-D stage B: e rel err %.3g, u rel err %.3g, fallbacks %d, mismatches %d" %
          (ee, ue, int(fo.sum()), int((fo != fg).sum())), flush=True)
    assert ee < 1e-6 and ue < 1e-6
    assert np.array_equal(fo, fg)
