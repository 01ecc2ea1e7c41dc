"""Independent pins of the oracle's 3-D streaming, 3-D charge pass,
E_DEP_FS and the EOS_TYPE 2 pressure, each shown to fail when one term of the
oracle is planted wrong (oracle.mutation).

3-D streaming (src/FreeStream.cpp:174-195, :250-273; src/LandauMatch.cpp:14-59,
:61-108).  Trilinear interpolation reproduces affine data exactly, so for an
affine initial density e0 = a + bx*ix + by*iy + be*ieta (dyadic coefficients,
exact in float) every kept sample's value is the same affine function at the
source point, and T^{mu nu} is a finite sum over rapidity nodes k and angles j
of closed-form terms.  The node factors are derived here from the
definitions, not copied from the oracle's expressions:
  y_k - eta = t_k = tan(pi/2 w_k),  w_k = -0.9975 + k 1.995/(N_rap - 1)
  Jacobian  = dy/dw = (pi/2) / cos^2(pi/2 w_k) * 1.995/(N_rap - 1)
  density factor R_k = N cosh^2(t_k) exp(-t_k^2 / 2 sigma^2),
                   N = 1 / (2 pi sqrt(pi/2) sigma (1 + exp(2 sigma^2)))
  source rapidity: sinh(eta' - y) = (tau/tau0) sinh(eta - y), so
                   eta' - eta = t_k - asinh(A sinh t_k), A = tau/tau0
  transverse shift D = tau cosh(y - eta') - tau0 cosh(y - eta)
                     = tau sqrt(1 + A^2 sinh^2 t_k) - tau0 cosh t_k
and a sample is kept when the floor of each grid index lies in [1, N-2].
Cells where a sample's index lies within rounding of a mask boundary are left
out of the comparison (the two computations may round to either side).

3-D charge pass (src/FreeStream.cpp:353-387, src/LandauMatch.cpp:110-143):
the same geometry with n0 = a + bx*ix + by*iy, whose rapidity factor
cosh(eta - r_k) exp(-(eta - r_k)^2 / sigma_B^2) on the uniform grid r_k is
interpolated linearly between eta planes; at DTAU = 0 on a fine grid,
n_B = n0 exp(-sigma_B^2/4)/sqrt(2) (the integral of cosh(u) exp(-u^2/s^2) is
s sqrt(pi) exp(s^2/4), against the norm's sqrt(2 pi) s exp(s^2/2)).

E_DEP_FS (src/FreeStream.cpp:389-417): on a two-valued profile the serial
float sums are exact, so DTAU = TAU_R (hbarc sum e^2 / sum e / E_R)^ALPHA in
closed form, for the library's and the oracle's parameter finalisation.

EOS_TYPE 2 (src/EquationOfState.cpp:20-69): a double polyval of the fit's
coefficients, read from the reference source by tests/golden/make_eos_table.py
into tests/golden/eos_wb_coefficients.json, and the fit's conformal plateau.
"""
import json
import os

import numpy as np
import pytest

import fsm_amd
import fsm_amd.workloads as W
import oracle as O

HB = float(np.float32(0.197326938))
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ------------------------------------------------------------ 3-D geometry --
def _params3d(case, baryon=False):
    if case == "a":
        kw = dict(DIM_X=47, DIM_Y=53, DIM_ETA=41, DIM_RAP=21, DIM_PHIP=16, DX=0.1, DY=0.09, DETA=0.1,
                  TAU0=0.1, DTAU=0.77, SIGMA=1.0)
        coef = (40.0, 0.25, 0.375, 0.5)
    else:  # x and y roles exchanged, another rapidity grid and time step
        kw = dict(DIM_X=53, DIM_Y=47, DIM_ETA=37, DIM_RAP=11, DIM_PHIP=24, DX=0.09, DY=0.1, DETA=0.12,
                  TAU0=0.2, DTAU=0.53, SIGMA=0.7)
        coef = (32.0, 0.375, 0.25, -0.25)
    p = W.base_params(**kw)
    p.IC_ENERGY = 3  # the input is the fm^-4 density itself (no unit conversion)
    if baryon:
        p.BARYON = 1
        p.SIGMA_B = 0.8
        p.DRAP = 0.3
    return p, coef


def _affine(q, coef, with_eta=True):
    a, bx, by, be = coef
    ix = np.arange(q.DIM_X, dtype=np.float64)[:, None, None]
    iy = np.arange(q.DIM_Y, dtype=np.float64)[None, :, None]
    ie = np.arange(q.DIM_ETA, dtype=np.float64)[None, None, :]
    v = a + bx * ix + by * iy + (be * ie if with_eta else 0.0 * ie)
    v32 = v.astype(np.float32)
    assert np.array_equal(v32.astype(np.float64), v)  # exact in float
    return v32.ravel()


def _f64(x):
    return float(np.float32(x))


def _nodes(q):
    """Per rapidity node: t_k, Jacobian, eta' - eta, D, cosh, tanh (double)."""
    nr = q.DIM_RAP
    w = -0.9975 + np.arange(nr) * (1.995 / (nr - 1))
    t = np.tan(np.pi / 2 * w)
    jac = (np.pi / 2) / np.cos(np.pi / 2 * w) ** 2 * (1.995 / (nr - 1))
    tau, tau0 = _f64(q.TAU), _f64(q.TAU0)
    A = tau / tau0
    with np.errstate(over="ignore", invalid="ignore"):
        s = A * np.sinh(t)
        deta = t - np.arcsinh(s)
        D = tau * np.sqrt(1.0 + s * s) - tau0 * np.cosh(t)
    return t, jac, deta, D, np.cosh(t), np.tanh(t)


def _hyper_rows(c, s, ch, th, tau):
    return np.array([1.0, c / ch, s / ch, th / tau, c * c / ch ** 2, c * s / ch ** 2, c * th / (tau * ch),
                     s * s / ch ** 2, s * th / (tau * ch), th * th / tau ** 2])


def _expected_stream(q, coef, eps, value_factor, rows=10, with_eta=True):
    """sum_k sum_j weight * keep * value, with the value of a kept sample
    = value_factor(k, ixn, iyn, ien) (arrays broadcast over the grid).
    Returns (T[rows][DIM], ambiguous[DIM])."""
    nx, ny, ne = q.DIM_X, q.DIM_Y, q.DIM_ETA
    t, jac, deta, D, ch, th = _nodes(q)
    tau = _f64(q.TAU)
    nphi = q.DIM_PHIP
    dphi = 2 * np.pi / nphi
    dx, dy, de = _f64(q.DX), _f64(q.DY), _f64(q.DETA)
    IX, IY, IE = np.arange(nx, dtype=np.float64), np.arange(ny, dtype=np.float64), np.arange(ne, dtype=np.float64)
    T = np.zeros((rows, nx, ny, ne))
    amb = np.zeros((nx, ny, ne), dtype=bool)

    def keep1(v, n):
        return (np.floor(v) >= 1) & (np.floor(v) <= n - 2)

    def amb1(v, n):
        return (np.abs(v - 1.0) < eps) | (np.abs(v - (n - 1.0)) < eps)

    for k in range(q.DIM_RAP):
        if not np.isfinite(D[k]) or not np.isfinite(deta[k]):
            continue
        ien = IE + deta[k] / de
        ke, ae = keep1(ien, ne), amb1(ien, ne)
        if not (ke.any() or ae.any()):
            continue
        for j in range(nphi):
            phi = j * 2 * np.pi / nphi
            c, s = np.cos(phi), np.sin(phi)
            ixn = IX - D[k] * c / dx
            iyn = IY - D[k] * s / dy
            kx, ky = keep1(ixn, nx), keep1(iyn, ny)
            ax, ay = amb1(ixn, nx), amb1(iyn, ny)
            amb |= (ax[:, None, None] & (ky | ay)[None, :, None] & (ke | ae)[None, None, :]) | \
                   ((kx | ax)[:, None, None] & ay[None, :, None] & (ke | ae)[None, None, :]) | \
                   ((kx | ax)[:, None, None] & (ky | ay)[None, :, None] & ae[None, None, :])
            if not (kx.any() and ky.any() and ke.any()):
                continue
            keep = kx[:, None, None] & ky[None, :, None] & ke[None, None, :]
            val = value_factor(k, ixn[:, None, None], iyn[None, :, None], ien[None, None, :]) * keep
            h = _hyper_rows(c, s, ch[k], th[k], tau)[:rows]
            wgt = dphi * jac[k] if value_factor.jacobian else dphi * value_factor.extra
            for m in range(rows):
                T[m] += (wgt * h[m]) * val
    return T.reshape(rows, -1), amb.ravel()


class _EnergyValue:
    jacobian = True

    def __init__(self, q, coef):
        t = _nodes(q)[0]
        sig = _f64(q.SIGMA)
        N = 1.0 / (2 * np.pi * np.sqrt(np.pi / 2) * sig * (1.0 + np.exp(2 * sig * sig)))
        with np.errstate(over="ignore", invalid="ignore"):
            self.R = N * np.cosh(t) ** 2 * np.exp(-t * t / (2 * sig * sig))
        self.a, self.bx, self.by, self.be = coef

    def __call__(self, k, ixn, iyn, ien):
        return self.R[k] * (self.a + self.bx * ixn + self.by * iyn + self.be * ien)


def _pin_3d_stream(case, precision):
    p, coef = _params3d(case)
    e = _affine(p, coef)
    q = O.params_finalize(p, e)
    eps, tol = (1e-9, 1e-10) if precision == "double" else (2e-3, 2e-5)
    T_exp, amb = _expected_stream(q, coef, eps, _EnergyValue(q, coef))
    r = O.run(q, e, precision, want_extra=False)
    keep = ~amb
    scale = np.abs(T_exp[0]).max()
    err = np.abs(r["T"][:, keep] - T_exp[:, keep]).max(axis=1) / scale
    assert keep.mean() > 0.9, keep.mean()
    assert np.all(err < tol), (case, precision, err)
    return err


@pytest.mark.parametrize("precision", ["float", "double"])
@pytest.mark.parametrize("case", ["a", "b"])
def test_3d_streaming_affine_closed_form(case, precision):
    err = _pin_3d_stream(case, precision)
    print(case, precision, "max rel err per component", np.array2string(err, precision=2))


@pytest.mark.parametrize("mut", ["trilinear_corner", "eta_new_sign", "shift_no_tau0", "jacobian_one_cos",
                                 "3d_over_tau", "hypertrig_tau", "eta_mask_low", "rap_factor_cosh"])
def test_3d_streaming_pin_catches_planted_mistakes(mut):
    with O.mutation(O.MUTATIONS[mut]):
        with pytest.raises(AssertionError):
            _pin_3d_stream("a", "double")


# ------------------------------------------------------------- 3-D charge --
class _ChargeValue:
    jacobian = False

    def __init__(self, q, coef, precision):
        sb = _f64(q.SIGMA_B)
        # SIGMA_B * SIGMA_B is a float product on the reference float path
        sb2 = float(np.float32(sb) * np.float32(sb)) if precision == "float" else sb * sb
        self.norm = 1.0 / (2 * np.pi * np.sqrt(2 * np.pi) * sb * np.exp(sb2 / 2))
        nr, drap = q.DIM_RAP, _f64(q.DRAP)
        self.r = -((nr - 1) / 2.0) * drap + np.arange(nr) * drap
        de = _f64(q.DETA)
        self.etas = -((q.DIM_ETA - 1) / 2.0) * de + np.arange(q.DIM_ETA) * de
        self.sb2 = sb2
        self.extra = drap  # J = DRAP * d_phip * sum (no Jacobian)
        self.a, self.bx, self.by, _ = coef

    def f(self, k, eta):
        d = eta - self.r[k]
        return np.cosh(d) * np.exp(-d * d / self.sb2)

    def __call__(self, k, ixn, iyn, ien):
        i0 = np.clip(np.floor(ien), 0, len(self.etas) - 2).astype(int)
        tt = ien - np.floor(ien)
        fe = (1 - tt) * self.f(k, self.etas[i0]) + tt * self.f(k, self.etas[i0 + 1])
        return self.norm * (self.a + self.bx * ixn + self.by * iyn) * fe


def _pin_3d_charge(case, precision):
    p, coef = _params3d(case, baryon=True)
    e = _affine(p, coef)
    nb = _affine(p, coef, with_eta=False)
    q = O.params_finalize(p, e)
    eps, tol = (1e-9, 1e-10) if precision == "double" else (2e-3, 2e-5)
    J_exp, amb = _expected_stream(q, coef, eps, _ChargeValue(q, coef, precision), rows=4)
    r = O.run(q, e, precision, nB=nb)
    J = r["V"] + r["nB"][None, :] * r["u"]  # J^mu = V^mu + n_B u^mu
    keep = ~amb
    scale = np.abs(J_exp[0]).max()
    err = np.abs(J[:, keep] - J_exp[:, keep]).max(axis=1) / scale
    assert keep.mean() > 0.9
    assert np.all(err < (tol if precision == "float" else 1e-9)), (case, precision, err)
    return err


@pytest.mark.parametrize("precision", ["float", "double"])
@pytest.mark.parametrize("case", ["a", "b"])
def test_3d_charge_affine_closed_form(case, precision):
    err = _pin_3d_charge(case, precision)
    print(case, precision, "J max rel err", np.array2string(err, precision=2))


@pytest.mark.parametrize("mut", ["charge_width", "charge_norm", "baryon_no_drap", "trilinear_corner",
                                 "shift_no_tau0"])
def test_3d_charge_pin_catches_planted_mistakes(mut):
    with O.mutation(O.MUTATIONS[mut]):
        with pytest.raises(AssertionError):
            _pin_3d_charge("a", "double")


def _pin_3d_charge_dtau0(precision):
    sb = 0.8
    p = W.base_params(DIM_X=9, DIM_Y=9, DIM_ETA=9, DIM_RAP=201, DIM_PHIP=8, DX=0.1, DY=0.1, DETA=0.1,
                      TAU0=0.5, DTAU=0.0, SIGMA=1.0)
    p.BARYON = 1
    p.SIGMA_B = sb
    p.DRAP = 0.05
    rng = np.random.default_rng(8)
    e = rng.uniform(1.0, 3.0, p.DIM).astype(np.float32)
    n0 = rng.uniform(0.1, 0.5, p.DIM).astype(np.float32)
    q = O.params_finalize(p, e)
    r = O.run(q, e, precision, nB=n0)
    inner = np.zeros((9, 9, 9), dtype=bool)
    inner[2:7, 2:7, 2:7] = True
    inner = inner.ravel()
    sb2 = float(np.float32(sb) * np.float32(sb)) if precision == "float" else _f64(sb) ** 2
    want = n0.astype(np.float64) * np.exp(-sb2 / 4) / np.sqrt(2)
    got = r["nB"]
    rel = np.abs(got - want)[inner].max() / want[inner].max()
    vx = max(np.abs(r["V"][1])[inner].max(), np.abs(r["V"][2])[inner].max()) / want[inner].max()
    centre = np.zeros((9, 9, 9), dtype=bool)
    centre[2:7, 2:7, 4] = True  # eta = 0: the odd tanh of the node against even rapidity factors
    veta = np.abs(r["V"][3])[centre.ravel()].max() / want[inner].max()
    tol = 1e-9 if precision == "double" else 1e-5  # float: 1608 float terms per cell
    assert rel < tol and vx < tol and veta < tol, (rel, vx, veta)
    return rel


@pytest.mark.parametrize("precision", ["float", "double"])
def test_3d_charge_dtau_zero_closed_form(precision):
    print(precision, "n_B rel err %.3g" % _pin_3d_charge_dtau0(precision))


@pytest.mark.parametrize("mut", ["charge_width", "charge_norm", "baryon_no_drap"])
def test_3d_charge_dtau_zero_pin_catches_planted_mistakes(mut):
    with O.mutation(O.MUTATIONS[mut]):
        with pytest.raises(AssertionError):
            _pin_3d_charge_dtau0("double")


# ---------------------------------------------------------------- E_DEP_FS --
def _edep_case(ic):
    p = W.base_params(DIM_X=40, DIM_Y=100, DIM_PHIP=8, DX=0.1, DY=0.08, TAU0=0.1, DTAU=0.5)
    p.E_DEP_FS = 1
    p.E_R, p.TAU_R, p.ALPHA = 0.25, 0.5, 0.4
    p.IC_ENERGY = ic
    v = np.full(4000, 0.5, dtype=np.float32)
    v[::4] = 2.0  # 1000 cells of 2, 3000 of 0.5
    if ic == 5:
        # GeV/fm^3 in; the reference sums the converted v/hbarc + 1e-7
        eps = v.astype(np.float64) / HB + 1e-7
    else:
        eps = v.astype(np.float64)
    e_T = HB * (eps ** 2).sum() / eps.sum()
    dtau = _f64(p.TAU_R) * (e_T / _f64(p.E_R)) ** _f64(p.ALPHA)
    return p, v, dtau


def _pin_edep(finalize, ic):
    p, v, dtau = _edep_case(ic)
    q = finalize(p, v)
    # ic 3: every partial float sum is exact; ic 5: 4000 rounded terms
    assert abs(q.DTAU - dtau) / dtau < (2e-6 if ic == 3 else 2e-5), (ic, q.DTAU, dtau)
    assert abs(q.TAU - (_f64(p.TAU0) + q.DTAU)) < 1e-6
    assert q.DIM_PHIP == max(8, int(np.ceil(2 * np.pi * q.DTAU / 0.08 - 1e-6)))
    return q.DTAU


@pytest.mark.parametrize("ic", [3, 5])
@pytest.mark.parametrize("who", ["oracle", "library"])
def test_energy_dependent_dtau_closed_form(who, ic):
    fin = O.params_finalize if who == "oracle" else fsm_amd.params_finalize
    _pin_edep(fin, ic)


@pytest.mark.parametrize("ic", [3, 5])
def test_energy_dependent_dtau_pin_catches_planted_mistake(ic):
    with O.mutation(O.MUTATIONS["edep_no_hbarc"]):
        with pytest.raises(AssertionError):
            _pin_edep(O.params_finalize, ic)


# ----------------------------------------------------------------- EOS 2 ---
def _eos_table():
    with open(os.path.join(GOLDEN, "eos_wb_coefficients.json")) as f:
        t = json.load(f)
    return np.array(t["a"]), np.array(t["b"])


def _pin_eos2(precision):
    a, b = _eos_table()
    # the float path forms e^12 in float, which overflows (p = inf/inf = NaN,
    # as in the reference's own float arithmetic) above e ~ 1600 fm^-4
    e = np.logspace(-3, 6, 181) if precision == "double" else np.logspace(-3, 3, 121)
    want = np.polynomial.polynomial.polyval(e, a) / np.polynomial.polynomial.polyval(e, b)
    got = O.pressure(e, 2, precision)
    tol = 1e-12 if precision == "double" else 2e-6
    rel = np.abs(got - want) / np.abs(want)
    assert rel.max() < tol, (precision, rel.max(), e[rel.argmax()])
    return rel.max()


@pytest.mark.parametrize("precision", ["float", "double"])
def test_eos2_against_polyval_of_the_fit(precision):
    print(precision, "EOS 2 max rel err %.3g" % _pin_eos2(precision))


def test_eos2_pin_catches_planted_mistake():
    with O.mutation(O.MUTATIONS["eos2_swap"]):
        with pytest.raises(AssertionError):
            _pin_eos2("double")


def test_eos2_fit_shape():
    """The fit's own limits: p(e -> 0) = a0/b0; p/e lies below the
    conformal 1/3 and reaches ~0.33 where a12 e^12 / b11 e^11 dominates
    (1e5 fm^-4); p rises with e."""
    a, b = _eos_table()
    p0 = O.pressure(np.array([1e-13]), 2, "double")[0]
    assert abs(p0 - a[0] / b[0]) < 1e-6 * abs(a[0] / b[0])
    e = np.logspace(-1, 5, 61)
    p = O.pressure(e, 2, "double")
    assert np.all(np.diff(p) > 0)
    assert np.all(p / e < 1.0 / 3.0) and np.all(p / e > 0.0)
    assert abs(p[-1] / e[-1] - a[12] / b[11]) < 0.01
    # conformal EoS and the lattice placeholder
    assert np.allclose(O.pressure(e, 1, "double"), e / 3.0, rtol=1e-15)
    assert np.all(O.pressure(e, 3, "double") == 0.0)
    # the reference float path's overflow of e^12
    assert np.isnan(O.pressure(np.array([2000.0]), 2, "float")[0])
