"""Input adapters and file formats of the reference, through libfsm.so's host
functions (no GPU work): the initial-condition dispatch
(src/FreestreamMilne.cpp:231-298, src/InitialConditions.cpp), the profile
readers (src/FileIO.cpp:133-164, src/InitialConditions.cpp:127-355), the
OUTPUTFORMAT=1 writers (src/FileIO.cpp:11-131) and the energy sums
(src/FreestreamMilne.cpp:531-559).

Pins: the C library's own rand() (ctypes, this container's glibc) for the
fluctuating generators' sequence; numpy restatements with the reference's
float/double mixing for the generators; Python's %g (the C99 rules that
std::ostream's default float output follows) for the writers; round trips
and hand-made files for the readers; closed forms for the sums."""
import ctypes
import os

import numpy as np
import pytest

import fsm_amd
import fsm_amd.workloads as W
import oracle as O


def _params(nx=7, ny=5, neta=3, **kw):
    p = W.base_params(DIM_X=nx, DIM_Y=ny, DIM_ETA=neta, DX=0.2, DY=0.3, DETA=0.25, **kw)
    return p


def _coords(p):
    # src/InitialConditions.cpp:34-37: float product minus a double half width, stored as float
    def c(n, d):
        i = np.arange(n, dtype=np.float32)
        return ((i * np.float32(d)).astype(np.float64) - np.float64(np.float32(n - 1)) / 2.0 *
                np.float64(np.float32(d))).astype(np.float32)

    return c(p.DIM_X, p.DX), c(p.DIM_Y, p.DY), c(p.DIM_ETA, p.DETA)


def _gauss_ref(p, bx, by, beta, rand=None):
    x, y, eta = _coords(p)
    X, Y, E = np.meshgrid(x, y, eta, indexing="ij")
    X, Y, E = X.ravel(), Y.ravel(), E.ravel()
    gx = np.exp(-(1.0 / np.float64(np.float32(bx))) * (X * X).astype(np.float64))
    gy = np.exp(-(1.0 / np.float64(np.float32(by))) * (Y * Y).astype(np.float64))
    ge = np.exp(-(1.0 / np.float64(np.float32(beta))) * (E * E).astype(np.float64))
    e0 = np.float64(500.0)
    if rand is not None:
        f = (rand.astype(np.float32) / np.float32(2147483647.0)).astype(np.float32)
        e0 = (np.float32(500.0) * f).astype(np.float64)
    return (((e0 * gx) * gy) * ge).astype(np.float32)


def _libc_rand(n, seed=1):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(seed)
    return np.array([libc.rand() for _ in range(n)], dtype=np.int64)


def test_rand_sequence_is_the_c_library_one():
    fsm_amd.ic_seed(1)
    ours = [fsm_amd.ic_rand() for _ in range(2000)]
    assert ours == list(_libc_rand(2000))
    fsm_amd.ic_seed(12345)
    assert [fsm_amd.ic_rand() for _ in range(100)] == list(_libc_rand(100, 12345))


def test_smooth_gaussians():
    p = _params(IC_ENERGY=1, IC_BARYON=1)
    e = fsm_amd.initial_condition(p, fsm_amd.IC_WHICH_ENERGY)
    assert np.array_equal(e, _gauss_ref(p, 15.0, 15.0, 15.0))
    nb = fsm_amd.initial_condition(p, fsm_amd.IC_WHICH_BARYON)
    assert np.array_equal(nb, _gauss_ref(p, 2.0, 3.0, 1.0))


def test_fluctuating_gaussians_follow_rand():
    p = _params(IC_ENERGY=2, IC_BARYON=2)
    D = p.DIM_X * p.DIM_Y * p.DIM_ETA
    fsm_amd.ic_seed(1)
    e = fsm_amd.initial_condition(p, fsm_amd.IC_WHICH_ENERGY)
    nb = fsm_amd.initial_condition(p, fsm_amd.IC_WHICH_BARYON)  # continues the sequence, as in one run
    r = _libc_rand(2 * D)
    assert np.array_equal(e, _gauss_ref(p, 15.0, 15.0, 15.0, r[:D]))
    assert np.array_equal(nb, _gauss_ref(p, 2.0, 3.0, 1.0, r[D:]))


def test_two_gaussians_are_zero_but_draw():
    p = _params(IC_ENERGY=6)
    D = p.DIM_X * p.DIM_Y * p.DIM_ETA
    fsm_amd.ic_seed(1)
    e = fsm_amd.initial_condition(p, fsm_amd.IC_WHICH_ENERGY)
    assert not e.any()
    assert fsm_amd.ic_rand() == _libc_rand(D + 1)[-1]


def test_dispatch_errors():
    p = _params(IC_ENERGY=5)
    with pytest.raises(fsm_amd.FsmError) as ex:
        fsm_amd.initial_condition(p, fsm_amd.IC_WHICH_ENERGY)
    assert ex.value.status == fsm_amd.FSM_ERR_IC
    p = _params(IC_ENERGY=3)
    with pytest.raises(fsm_amd.FsmError) as ex:
        fsm_amd.initial_condition(p, fsm_amd.IC_WHICH_ENERGY, "/nonexistent_dir_xyz")
    assert ex.value.status == fsm_amd.FSM_ERR_FILE
    p = _params(IC_BARYON=7)
    with pytest.raises(fsm_amd.FsmError):
        fsm_amd.initial_condition(p, fsm_amd.IC_WHICH_BARYON)


def test_xyev_reader_round_trip(tmp_path):
    p = _params(IC_ENERGY=3)
    x, y, eta = _coords(p)
    rng = np.random.default_rng(3)
    vals = rng.uniform(0, 10, size=(p.DIM_X, p.DIM_Y, p.DIM_ETA)).astype(np.float32)
    lines = []
    for ix in range(p.DIM_X):
        for iy in range(p.DIM_Y):
            for ie in range(p.DIM_ETA):
                lines.append("%.6f %.6f %.6f %.9g" % (x[ix], y[iy], eta[ie], vals[ix, iy, ie]))
    lines.append("100.0 0.0 0.0 5.0")  # outside the grid: skipped
    rng.shuffle(lines)
    (tmp_path / "e.dat").write_text("\n".join(lines) + "\n")
    got = fsm_amd.initial_condition(p, fsm_amd.IC_WHICH_ENERGY, str(tmp_path))
    assert np.array_equal(got, vals.ravel())


def test_cpuvh_reader(tmp_path):
    p = _params(IC_ENERGY=4)
    vals = np.arange(p.DIM_X * p.DIM_Y * p.DIM_ETA, dtype=np.float32).reshape(p.DIM_X, p.DIM_Y, p.DIM_ETA) * 0.5
    rows = []
    for ie in range(p.DIM_ETA):  # eta, y, x order
        for iy in range(p.DIM_Y):
            for ix in range(p.DIM_X):
                rows.append("0 0 0 %.9g" % vals[ix, iy, ie])
    (tmp_path / "e.dat").write_text("\n".join(rows) + "\n")
    got = fsm_amd.initial_condition(p, fsm_amd.IC_WHICH_ENERGY, str(tmp_path))
    assert np.array_equal(got, (vals + np.float32(1e-7)).ravel())


def _envelope(p):
    _, _, eta = _coords(p)
    a = (np.abs(eta) - np.float32(p.ETA_FLAT)).astype(np.float32)
    arg = ((-1.0) * a.astype(np.float64) * a / (2.0 * np.float64(np.float32(p.ETA_WIDTH)) ** 2)).astype(np.float32)
    arg = arg * np.where(a < 0, np.float32(0), np.float32(1))
    return np.exp(arg.astype(np.float32)).astype(np.float32)


def test_block_readers(tmp_path):
    p = _params(neta=9)
    p.ETA_FLAT, p.ETA_WIDTH = 0.4, 0.3
    blk = np.arange(p.DIM_X * p.DIM_Y, dtype=np.float32).reshape(p.DIM_X, p.DIM_Y) + 1.0
    env = _envelope(p)
    want = (blk[:, :, None].astype(np.float64) * env[None, None, :].astype(np.float64) + 1e-3).astype(np.float32)
    # TRENTO: 12 header lines, then y rows of x values
    txt = "".join("# header %d\n" % k for k in range(12))
    txt += "\n".join(" ".join("%.9g" % blk[ix, iy] for ix in range(p.DIM_X)) for iy in range(p.DIM_Y)) + "\n"
    (tmp_path / "t.dat").write_text(txt)
    got = fsm_amd.read_profile(p, fsm_amd.PROFILE_TRENTO_BLOCK, str(tmp_path / "t.dat"))
    np.testing.assert_allclose(got, want.ravel(), rtol=4e-7)  # float vs double rounding of the envelope
    # superMC: x rows of y values, no header
    txt = "\n".join(" ".join("%.9g" % blk[ix, iy] for iy in range(p.DIM_Y)) for ix in range(p.DIM_X)) + "\n"
    (tmp_path / "s.dat").write_text(txt)
    got = fsm_amd.read_profile(p, fsm_amd.PROFILE_SUPERMC_BLOCK, str(tmp_path / "s.dat"))
    np.testing.assert_allclose(got, want.ravel(), rtol=4e-7)  # float vs double rounding of the envelope
    # TRENTO 3-D: every cell in cell order, + 1e-3
    v3 = np.linspace(0, 3, p.DIM_X * p.DIM_Y * p.DIM_ETA).astype(np.float32)
    (tmp_path / "t3.dat").write_text(" ".join("%.9g" % v for v in v3))
    got = fsm_amd.read_profile(p, fsm_amd.PROFILE_TRENTO3D_BLOCK, str(tmp_path / "t3.dat"))
    assert np.array_equal(got, v3 + np.float32(1e-3))


def test_reference_profile_files():
    """The reference's own example profiles (initial_profiles/e.dat, a 131^2
    TRENTO event with a 12-line header, rows along y; initial_profiles/
    ed_event_1_block.dat, a 51^2 superMC block, rows along x), committed as
    tests/golden/trento_e.dat and supermc_ed_event_1_block.dat, read exactly:
    value + 1e-3f on a 2-D grid (the eta envelope is 1 at eta = 0)."""
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    p = _params()
    p.DIM_X = p.DIM_Y = 131
    p.DIM_ETA = 1
    got = fsm_amd.read_profile(p, fsm_amd.PROFILE_TRENTO_BLOCK, os.path.join(gold, "trento_e.dat"))
    data = np.loadtxt(os.path.join(gold, "trento_e.dat"), comments="#", dtype=np.float64)
    assert data.shape == (131, 131)
    tol = np.float64(np.float32(1e-3))  # the reference's float lower_tolerance, added in double
    want = (data.T.astype(np.float32).astype(np.float64) + tol).astype(np.float32)  # [ix, iy]
    assert np.array_equal(got, want.ravel())
    p.DIM_X = p.DIM_Y = 51
    got = fsm_amd.read_profile(p, fsm_amd.PROFILE_SUPERMC_BLOCK, os.path.join(gold, "supermc_ed_event_1_block.dat"))
    data = np.loadtxt(os.path.join(gold, "supermc_ed_event_1_block.dat"), dtype=np.float64)
    assert data.shape == (51, 51)
    want = (data.astype(np.float32).astype(np.float64) + tol).astype(np.float32)
    assert np.array_equal(got, want.ravel())


def _snapped(n, d):
    i = np.arange(n, dtype=np.float32)
    x = ((i * np.float32(d)).astype(np.float64) - (np.float64(np.float32(n - 1)) / 2.0 * np.float64(np.float32(d))))
    x = x.astype(np.float32)
    return (np.float32(d) * np.round(x / np.float32(d)).astype(np.float32)).astype(np.float32)


def test_scalar_writer_format(tmp_path):
    p = _params()
    D = p.DIM_X * p.DIM_Y * p.DIM_ETA
    var = (np.random.default_rng(1).standard_normal(D) * 1e3).astype(np.float32)
    var[0], var[1] = 0.0, 1.5e-7
    f = tmp_path / "e.dat"
    fsm_amd.write_scalar(p, str(f), var)
    x, y, eta = _snapped(p.DIM_X, p.DX), _snapped(p.DIM_Y, p.DY), _snapped(p.DIM_ETA, p.DETA)
    want = []
    for ix in range(p.DIM_X):
        for iy in range(p.DIM_Y):
            for ie in range(p.DIM_ETA):
                is_ = (p.DIM_Y * p.DIM_ETA) * ix + p.DIM_ETA * iy + ie
                want.append("%g %g %g %g" % (x[ix], y[iy], eta[ie], var[is_]))
    assert f.read_text() == "\n".join(want) + "\n"


def test_projection_writer_format(tmp_path):
    p = _params(neta=5)
    D = p.DIM_X * p.DIM_Y * p.DIM_ETA
    var = np.arange(D, dtype=np.float32) / 7
    f = tmp_path / "p.dat"
    fsm_amd.write_projection(p, str(f), var)
    ie = (p.DIM_ETA - 1) // 2
    want = ""
    for iy in range(p.DIM_Y):
        want += "".join("%g " % var[(p.DIM_Y * p.DIM_ETA) * ix + p.DIM_ETA * iy + ie] for ix in range(p.DIM_X)) + "\n"
    assert f.read_text() == want


def test_energy_sums():
    p = _params(neta=1)
    p.TAU = np.float32(p.TAU0) + np.float32(p.DTAU)
    D = p.DIM_X * p.DIM_Y
    T00 = np.ones(D, np.float32)
    e = np.full(D, 100.0, np.float32)
    e[:3] = 0.1  # below E_FREEZE / hbarc
    tot, ins = fsm_amd.energy_sums(p, T00, e)
    vol = np.float32(np.float32(p.TAU) * np.float32(p.DX)) * np.float32(p.DY)
    assert tot == pytest.approx(float(np.float32(D) * vol), rel=1e-6)
    assert ins == pytest.approx(float(np.float32(D - 3) * vol), rel=1e-6)
    rng = np.random.default_rng(4)
    T00 = rng.uniform(0, 3, D).astype(np.float32)
    e = rng.uniform(0, 20, D).astype(np.float32)
    assert fsm_amd.energy_sums(p, T00, e) == O.energy_sums(p, T00, e)


def test_energy_inside_file(tmp_path):
    f = tmp_path / "energy_inside_switch_surf.dat"
    fsm_amd.lib().fsm_write_energy_inside(str(f).encode(), ctypes.c_float(12.5))
    assert f.read_text() == "%g\n" % np.float32(np.float32(12.5) * np.float32(0.197326938))


def test_oracle_diagnostics_known_answers():
    p = _params(nx=9, ny=9, neta=1)
    p.TAU = 0.6
    D = p.DIM_X * p.DIM_Y
    x, y, _ = _coords(p)
    X, Y = np.meshgrid(x, y, indexing="ij")
    e = np.full(D, 7.0, np.float32)
    pr = (e / 3).astype(np.float32)
    Pi = np.full(D, 0.25, np.float32)
    a = 0.05
    u = np.zeros((4, D), np.float32)
    u[0] = 1.0
    u[1] = (-a * Y).ravel()
    u[2] = (a * X).ravel()
    pi = np.zeros((10, D), np.float32)
    pi[4], pi[7], pi[9] = 0.3, -0.1, -0.2 / p.TAU ** 2  # xx, yy, eta-eta: pi^xx + pi^yy + tau^2 pi^nn = 0
    d = O.diagnostics(p, e, u, pr, Pi, pi)
    np.testing.assert_allclose(d["R_Pi_Inv"], 0.25 / pr, rtol=1e-6)
    num = 0.3 ** 2 + 0.1 ** 2 + 0.2 ** 2
    np.testing.assert_allclose(d["R_pimunu_Inv"], np.sqrt(num) / np.sqrt(7.0 ** 2 + 3 * (7.0 / 3) ** 2), rtol=1e-5)
    np.testing.assert_allclose(d["pi_mu_mu"], 0.0, atol=1e-6)
    # rest-frame pi with u = (1, ...): pi^{mu nu} u_nu has x, y parts from the flow only
    np.testing.assert_allclose(d["pi_dot_u"][0], 0.0, atol=1e-7)
    # rigid rotation u_x = a y, u_y = -a x at constant T: w_xy = a / T on the interior
    T = (7.0 / 13.8997) ** 0.25
    w = d["omega"][3].reshape(p.DIM_X, p.DIM_Y)
    np.testing.assert_allclose(w[1:-1, 1:-1], a / T, rtol=1e-4)
    assert not w[0].any() and not w[:, 0].any()
    np.testing.assert_allclose(d["beta"][0].reshape(9, 9)[1:-1, 1:-1], 1 / T, rtol=1e-6)
