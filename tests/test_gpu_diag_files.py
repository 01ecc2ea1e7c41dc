"""GPU tests of the diagnostics kernel, the whole reference run with its files
(fsm_run_files) and the non-vector initial conditions, all through
libfsm.so's C ABI, against the oracle on the same inputs.

Diagnostics: the device kernel (src/HydroValidity.cpp:8-101,
src/LandauMatch.cpp:395-477) on the oracle's own float hydro fields must give
the oracle's values (same inputs, same float expressions; tolerance 1e-6
relative for the algebraic fields, 1e-5 for beta and omega, whose T(e) uses
powf).  The run with files: every output/<name>.dat the reference writes, with
values matching the oracle's float path to the parity tolerances."""
import os

import numpy as np
import pytest

import fsm_amd
import fsm_amd.workloads as W
import oracle as O
from parity import HBARC, TOL, output_errs, well_conditioned

pytestmark = pytest.mark.gpu


def _fields(out):
    """Float hydro fields from the 17 double planes: (double)x * hbarc is
    exact in double, so x = d / hbarc exactly."""
    hb = np.float64(HBARC)
    e = (out[0] / hb).astype(np.float32)
    p = (out[1] / hb).astype(np.float32)
    u = out[2:6].astype(np.float32)
    pi = (out[6:16] / hb).astype(np.float32)
    Pi = (out[16] / hb).astype(np.float32)
    return e, u, p, Pi, pi


@pytest.mark.parametrize("name", ["C0", "small3d"])
def test_diagnostics_kernel_matches_oracle(gpu_ctx, name):
    import torch

    p, e_in = W.config(name) if name == "C0" else W.small_3d()
    q = fsm_amd.params_finalize(p, e_in)
    ref = O.run(q, e_in, "float")
    e, u, pr, Pi, pi = _fields(ref["out"])
    want = O.diagnostics(q, e, u, pr, Pi, pi)
    D = q.DIM
    dev = [torch.from_numpy(np.ascontiguousarray(a).ravel()).cuda() for a in (e, u, pr, Pi, pi)]
    outs = torch.full((17, D), 7.0, dtype=torch.float32, device="cuda")  # every entry must be written
    o = fsm_amd.DiagOut()
    o.R_Pi_Inv = outs[0].data_ptr()
    o.R_pimunu_Inv = outs[1].data_ptr()
    for c in range(4):
        o.pi_dot_u[c] = outs[2 + c].data_ptr()
        o.beta[c] = outs[7 + c].data_ptr()
    o.pi_mu_mu = outs[6].data_ptr()
    for c in range(6):
        o.omega[c] = outs[11 + c].data_ptr()
    gpu_ctx.diagnostics_device(q, *[d.data_ptr() for d in (dev[0], dev[1], dev[2], dev[3], dev[4])], o)
    torch.cuda.synchronize()
    got = outs.cpu().numpy()
    print("case", name, "DIM", D, flush=True)

    def close(a, b, rtol, what):
        a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
        scale = max(np.abs(b).max(), 1e-30)
        err = np.abs(a - b).max() / scale
        print("  %-14s max err / max |ref| = %.3g" % (what, err))
        assert err <= rtol, what

    close(got[0], want["R_Pi_Inv"], 1e-6, "R_Pi_Inv")
    close(got[1], want["R_pimunu_Inv"], 1e-6, "R_pimunu_Inv")
    close(got[2:6], want["pi_dot_u"], 1e-6, "pi_dot_u")
    close(got[6], want["pi_mu_mu"], 1e-6, "pi_mu_mu")
    close(got[7:11], want["beta"], 1e-5, "beta")
    close(got[11:17], want["omega"], 1e-5, "omega")
    # entries the reference's loop never visits (2-D index DIM_Y*ix + iy over
    # the transverse interior) are exactly 0
    visited = np.zeros(D, dtype=bool)
    ix, iy = np.meshgrid(np.arange(1, q.DIM_X - 1), np.arange(1, q.DIM_Y - 1), indexing="ij")
    visited[(q.DIM_Y * ix + iy).ravel()] = True
    assert not got[7:17][:, ~visited].any()


EXPECTED_FILES = [
    "initial_e", "initial_e_projection", "energy_inside_switch_surf", "R_Pi_Inv_projection",
    "R_pimunu_Inv_projection", "pi_dot_u_tau_projection", "pi_dot_u_x_projection", "pi_dot_u_y_projection",
    "pi_dot_u_eta_projection", "pi_mu_mu_projection", "e", "p", "bulk_PI", "e_projection", "p_projection",
    "bulk_PI_projection", "beta_x", "beta_x_projection", "beta_y", "beta_y_projection", "w_xy", "w_xy_projection",
] + [n + s for n in ("ut", "ux", "uy", "un", "pitt", "pitx", "pity", "pitn", "pixx", "pixy", "pixn", "piyy",
                     "piyn", "pinn") for s in ("", "_projection")]


def test_run_files_writes_the_reference_set(gpu_ctx, tmp_path):
    p, e_in = W.config("C0")
    p = p.copy()
    p.OUTPUTFORMAT = 1  # vectors stay zero, files are written anyway
    outs, q, rep = gpu_ctx.run_files(p, e_in, None, str(tmp_path))
    names = sorted(f[:-4] for f in os.listdir(tmp_path))
    assert names == sorted(EXPECTED_FILES)
    assert all(not outs[k].any() for k in outs)
    ref = O.run(q, e_in, "float")
    e_ref = ref["e"].astype(np.float32)  # the 17 vectors are zero with OUTPUTFORMAT = 1
    cols = np.loadtxt(tmp_path / "e.dat")
    assert cols.shape == (q.DIM, 4)
    m = e_ref > 1e-3 * e_ref.max()
    err = np.abs(cols[:, 3] - e_ref)[m] / e_ref[m]
    print("e.dat vs oracle e: max rel err %.3g" % err.max())
    assert err.max() < TOL
    tot, ins = O.energy_sums(q, ref["T"][0].astype(np.float32), e_ref)
    print("energy after %.6g (oracle %.6g), inside %.6g (oracle %.6g)" % (rep["total_energy_after"], tot,
                                                                            rep["energy_inside"], ins))
    assert rep["total_energy_after"] == pytest.approx(tot, rel=1e-5)
    assert rep["energy_inside"] == pytest.approx(ins, rel=1e-3)  # cells at the E_FREEZE threshold may flip
    assert rep["tau_landau_match"] == pytest.approx(float(q.TAU))
    proj = np.loadtxt(tmp_path / "e_projection.dat")
    assert proj.shape == (q.DIM_Y, q.DIM_X)
    inside_txt = (tmp_path / "energy_inside_switch_surf.dat").read_text()
    assert float(inside_txt) == pytest.approx(rep["energy_inside"] * HBARC, rel=1e-5)


@pytest.mark.parametrize("case", ["gauss2d", "gauss3d"])
def test_non_vector_initial_condition(gpu_ctx, case):
    """IC_ENERGY = 1 (the smooth elliptical Gaussian, fm^-4, no GeV
    conversion): GPU step against the oracle's float path on the same
    profile."""
    if case == "gauss2d":
        p = W.base_params(DIM_X=61, DIM_Y=61, DIM_PHIP=40, DX=0.2, DY=0.2, TAU0=0.2, DTAU=1.0, IC_ENERGY=1)
    else:
        p = W.base_params(DIM_X=33, DIM_Y=33, DIM_ETA=17, DIM_RAP=7, DIM_PHIP=16, DX=0.25, DY=0.25, DETA=0.2,
                          TAU0=0.2, DTAU=0.6, SIGMA=1.0, IC_ENERGY=1)
    e = fsm_amd.initial_condition(p, fsm_amd.IC_WHICH_ENERGY)
    q = fsm_amd.params_finalize(p, e)
    outs, q2, _ = gpu_ctx.run(p, e)
    got = np.stack([outs[k] for k in fsm_amd.OUTPUT_NAMES])
    ref = O.run(O.params_finalize(p, e), e, "float")
    T00max = np.abs(ref["T"][0]).max()
    errs = output_errs(got, ref["out"], T00max, well_conditioned(ref["T"], q.TAU))
    print(case, {k: "%.3g" % v for k, v in errs.items()})
    assert max(errs[k] for k in ("e", "p", "u", "pi", "Pi")) < TOL


def test_python_mirror_runs_like_the_reference(gpu_ctx, tmp_path, monkeypatch):
    """fsm_amd.FreestreamMilne with a freestream_input selecting the smooth
    Gaussian (IC_ENERGY = 1): reads the file positionally, makes the profile,
    writes output/*.dat and fills the 17 vectors."""
    monkeypatch.chdir(tmp_path)
    (tmp_path / "output").mkdir()
    vals = [("OUTPUTFORMAT", 2), ("BARYON", 0), ("IC_ENERGY", 1), ("IC_BARYON", 1), ("ETA_WIDTH", 0.5),
            ("ETA_FLAT", 0.5), ("SIGMA", 1.0), ("SIGMA_B", 1.0), ("DIM_X", 41), ("DIM_Y", 41), ("DIM_ETA", 1),
            ("DIM_RAP", 1), ("DIM_PHIP", 30), ("DX", 0.25), ("DY", 0.25), ("DETA", 0.1), ("DRAP", 0.2),
            ("DTAU", 0.5), ("TAU0", 0.2), ("EOS_TYPE", 1), ("E_FREEZE", 1.7), ("VISCOUS_MATCHING", 1),
            ("E_DEP_FS", 0), ("E_R", 0.0), ("TAU_R", 0.0), ("ALPHA", 0.0)]
    (tmp_path / "freestream_input").write_text("".join("%s\t%s\n" % kv for kv in vals))
    fs = fsm_amd.FreestreamMilne(context=gpu_ctx)
    assert fs.run_freestream_milne() == 0
    e_out = fs.output_to_vectors()[0]
    assert e_out.size == 41 * 41 and e_out.max() > 0
    assert (tmp_path / "output" / "e.dat").exists() and (tmp_path / "output" / "w_xy.dat").exists()
    assert fs.tau_LandauMatch == pytest.approx(0.7)
    assert 0 < fs.report["energy_inside"] <= fs.report["total_energy_after"]
