"""The 4x4 Landau frame on the GPU, both paths, every cell against the
oracle's FP64 eigen-solve (the restatement of solveEigenSystem,
src/LandauMatch.cpp:145-331) of the same float tensor, with no conditioning
exclusion, and the reference's fallback decision (e = 1e-3, u = (1,0,0,0))
on every cell:

  * synthetic tensors with a known frame: rest-frame diag(e, P1, P2, P3),
    rotated and boosted up to gamma 20, from nearly null (two pressures near
    zero) to isotropic; also against the exact e and u, within what the
    float rounding of T allows (kappa times the float epsilon);
  * the oracle's T^{mu nu} of 3-D events (eta-major on the run path,
    k_landau3d; reference layout through fsm_landau_device, k_landau_one).

Both paths solve every cell in FP64 (the characteristic quartic, Jacobi
sweeps where it does not settle); the bar is the float rounding of the
results."""
import numpy as np
import pytest

import fsm_amd
import fsm_amd.workloads as W
import oracle as O
from test_oracle_landau import _boost

pytestmark = pytest.mark.gpu

TAU0, DTAU = 0.1, 0.8
TAU = float(np.float32(TAU0) + np.float32(DTAU))
BAR = {"fp32": 1.0e-6, "mixed": 1.0e-6}  # e and u against FP64 on the same float tensor


def _ctx(path):
    ctx = fsm_amd.Context(0)
    ctx.set_path(fsm_amd.PATH_MIXED if path == "mixed" else fsm_amd.PATH_FP32)
    return ctx


def _known_frames(n, seed):
    """n tensors (Milne components, float32) with known e and u."""
    rng = np.random.default_rng(seed)
    T = np.empty((n, 10))
    e = np.empty(n)
    u = np.empty((n, 4))
    for k in range(n):
        ek = 10.0 ** rng.uniform(-3, 3)
        kind = k % 4
        if kind == 0:  # conformal, isotropic
            P = np.full(3, ek / 3)
        elif kind == 1:  # anisotropic
            P = rng.uniform(0.0, 1.0, 3) * ek / 3
        elif kind == 2:  # nearly null: free streaming along one axis
            P = np.array([ek * rng.uniform(0.3, 1.0), ek * 1e-4, ek * 1e-5])
        else:  # no longitudinal pressure (2-D-like), full frame
            P = np.array([ek / 2, ek / 2 * rng.uniform(0.2, 1.0), 0.0])
        R = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        Srest = np.zeros((4, 4))
        Srest[0, 0] = ek
        Srest[1:, 1:] = R @ np.diag(P) @ R.T
        gamma = 1.0 + (20.0 - 1.0) * rng.uniform() ** 3
        nv = rng.normal(size=3)
        nv /= np.linalg.norm(nv)
        uk = np.concatenate([[gamma], np.sqrt(gamma * gamma - 1.0) * nv])
        L = _boost(uk)
        A = L @ Srest @ L.T
        Sm = np.diag([1.0, 1.0, 1.0, 1.0 / TAU])
        Tm = Sm @ A @ Sm
        T[k] = [Tm[0, 0], Tm[0, 1], Tm[0, 2], Tm[0, 3], Tm[1, 1], Tm[1, 2], Tm[1, 3], Tm[2, 2], Tm[2, 3], Tm[3, 3]]
        e[k] = ek
        u[k] = [uk[0], uk[1], uk[2], uk[3] / TAU]
    return T.astype(np.float32), e, u


def _kappa(T, e, u):
    """Sensitivity of e to a relative change of every component (orthonormal
    frame): sum |u_mu| |T^{mu nu}| |u_nu| / e."""
    a = np.abs(u) * np.array([1.0, 1.0, 1.0, TAU])
    t = np.abs(T.astype(np.float64)) * np.array([1, 1, 1, TAU, 1, 1, TAU, 1, TAU, TAU * TAU])
    S = (a[:, 0] ** 2 * t[:, 0] + a[:, 1] ** 2 * t[:, 4] + a[:, 2] ** 2 * t[:, 7] + a[:, 3] ** 2 * t[:, 9] +
         2 * (a[:, 0] * (a[:, 1] * t[:, 1] + a[:, 2] * t[:, 2] + a[:, 3] * t[:, 3]) +
              a[:, 1] * (a[:, 2] * t[:, 5] + a[:, 3] * t[:, 6]) + a[:, 2] * a[:, 3] * t[:, 8]))
    return S / e


def _gpu_landau(ctx, T):
    """fsm_landau_device on cells given as (n, 10) float32 in the reference
    layout of a 3-D grid (the 4x4 frame)."""
    import torch

    n = T.shape[0]
    assert n % 64 == 0
    p = W.base_params(DIM_X=n // 64, DIM_Y=8, DIM_ETA=8, DIM_RAP=5, DIM_PHIP=8, TAU0=TAU0, DTAU=DTAU)
    q = fsm_amd.params_finalize(p, np.ones(n, dtype=np.float32))
    assert q.DIM == n and q.DIM_ETA > 1 and np.float32(q.TAU) == np.float32(TAU)
    d_T = torch.from_numpy(np.ascontiguousarray(T.T.ravel())).cuda()
    d_e = torch.empty(n, dtype=torch.float32, device="cuda")
    d_u = torch.empty(4 * n, dtype=torch.float32, device="cuda")
    ctx.landau_device(q, d_T.data_ptr(), d_e.data_ptr(), d_u.data_ptr(), stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    return d_e.cpu().numpy().astype(np.float64), d_u.cpu().numpy().reshape(4, n).T.astype(np.float64)


def _fallback(e, u):
    """Cells given the reference's fallback frame (float e = 1e-3 for the GPU,
    the double 1e-3 for the oracle)."""
    return (e.astype(np.float32) == np.float32(1e-3)) & (u[:, 0] == 1.0) & np.all(u[:, 1:] == 0.0, axis=1)


@pytest.mark.parametrize("path", ["fp32", "mixed"])
def test_landau4_known_frames(path):
    n_ok = 64 * 8 * 6
    T, e_x, u_x = _known_frames(n_ok, seed=41)
    # cells the reference rejects: zero tensors, negated tensors (e < 0)
    T = np.concatenate([T, np.zeros((64, 10), np.float32), -T[:64 * 7]])
    n = T.shape[0]
    ctx = _ctx(path)
    eg, ug = _gpu_landau(ctx, T)
    ctx.close()
    eo, uo = O.landau(T.astype(np.float64), np.float32(TAU))
    fo, fg = _fallback(eo, uo), _fallback(eg, ug)
    assert not fo[:n_ok].any() and fo[n_ok:].all()
    assert np.array_equal(fo, fg)
    ee = np.abs(eg - eo)[:n_ok] / eo[:n_ok]
    ue = np.abs(ug - uo)[:n_ok].max(axis=1) / uo[:n_ok, 0]
    kap = _kappa(T[:n_ok], e_x, u_x)
    # the exact frame, within the float rounding of T
    ex = np.abs(eo[:n_ok] - e_x) / e_x
    ux = np.abs(uo[:n_ok] - u_x).max(axis=1) / u_x[:, 0]
    print(path, "known frames: kappa up to %.3g; GPU vs FP64 e %.3g u %.3g; FP64 vs exact e %.3g u %.3g"
          % (kap.max(), ee.max(), ue.max(), (ex / (1 + kap)).max(), (ux / (1 + kap)).max()), flush=True)
    assert kap.max() > 1e3  # the set reaches ill-conditioned cells
    assert ee.max() < BAR[path] and ue.max() < BAR[path]
    assert np.all(ex <= 1e-6 * (1.0 + kap)) and np.all(ux <= 1e-6 * (1.0 + kap))


def _case3d(name):
    if name == "small3d":
        return W.small_3d(), "tan"
    if name == "small3d_int":  # DTAU/DX integer: samples on grid lines, nearly null single-beam cells
        return W.small_3d(nx=33, neta=17, nrap=7, nphi=16, dtau=0.5, seed=9), "tan"
    if name == "rect3d":
        from test_gpu_mixed import _case as mixed_case

        return mixed_case("rect3d"), "tan"
    if name == "c2_71_gh":  # C2-style event on a 71^2 x 41 grid, Gauss-Hermite rapidity rule
        return W.c2_like(41, nx=71), "gauss_hermite"
    raise ValueError(name)


@pytest.mark.parametrize("path", ["fp32", "mixed"])
@pytest.mark.parametrize("name", ["small3d", "small3d_int", "rect3d", "c2_71_gh"])
def test_landau4_on_oracle_tensors(name, path):
    """Stage B of 3-D events on identical input: the oracle's float T through
    fsm_landau_device, every cell with e > 1e-3 e_max."""
    (p, e), rule = _case3d(name)
    q = O.params_finalize(p, e)
    T = O.run(q, e, "float", want_extra=False, rapidity=rule)["T"].astype(np.float32).T
    pad = (-T.shape[0]) % 512
    T = np.concatenate([T, np.zeros((pad, 10), np.float32)])
    ctx = _ctx(path)
    eg, ug = _gpu_landau(ctx, T)
    ctx.close()
    eo, uo = O.landau(T.astype(np.float64), np.float32(TAU))
    m = eo > 1e-3 * eo.max()
    ee = (np.abs(eg - eo)[m] / eo[m]).max()
    ue = (np.abs(ug - uo)[m].max(axis=1) / uo[m, 0]).max()
    mism = int((_fallback(eo, uo) != _fallback(eg, ug)).sum())
    print(name, path, "3-D stage B: cells %d, e rel err %.3g, u rel err %.3g, fallback mismatches %d"
          % (int(m.sum()), ee, ue, mism), flush=True)
    assert ee < BAR[path] and ue < BAR[path]
    assert mism == 0


@pytest.mark.parametrize("path", ["fp32", "mixed"])
@pytest.mark.parametrize("name", ["small3d", "small3d_int", "rect3d"])
def test_landau4_frame_on_gpu_tensor(name, path):
    """The run path's 3-D Landau stage (eta-major tiles) on the GPU's own
    T^{mu nu}: e and u of every cell with e > 1e-3 e_max against the oracle's
    FP64 eigen-solve of the same tensor, and the reference's fallback
    decision on every cell (no conditioning exclusion)."""
    import torch

    (p, e), _ = _case3d(name)
    ctx = _ctx(path)
    q = fsm_amd.params_finalize(p, e)
    D = q.DIM
    d_e = torch.from_numpy(np.ascontiguousarray(e, dtype=np.float32)).cuda()
    d_T = torch.empty(10 * D, dtype=torch.float32, device="cuda")
    d_out = torch.empty(17 * D, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream()
    ctx.stream_tmunu_device(q, d_e.data_ptr(), d_T.data_ptr(), s)
    ctx.run_device(q, d_e.data_ptr(), d_out.data_ptr(), s)
    torch.cuda.synchronize()
    ctx.close()
    T = d_T.cpu().numpy().reshape(10, D)
    out = d_out.cpu().numpy().reshape(17, D)
    hb = np.float64(np.float32(0.197326938))
    eg = out[0] / hb  # (double)e * hbarc is exact, so this is the float e
    ug = out[2:6].T
    eo, uo = O.landau(T.astype(np.float64).T, np.float32(q.TAU))
    m = eo > 1e-3 * eo.max()
    ee = (np.abs(eg - eo)[m] / eo[m]).max()
    ue = (np.abs(ug - uo)[m].max(axis=1) / uo[m, 0]).max()
    fo, fg = _fallback(eo, uo), _fallback(eg, ug)
    print(name, path, "3-D frame on the GPU's T: cells %d, e rel err %.3g, u rel err %.3g, fallbacks %d, mismatches %d"
          % (int(m.sum()), ee, ue, int(fo.sum()), int((fo != fg).sum())), flush=True)
    assert ee < BAR[path] and ue < BAR[path]
    assert np.array_equal(fo, fg)
