"""Gauss-Hermite rapidity rule (FSM_RAPIDITY_GAUSS_HERMITE), CPU checks.

The reference ships 21-, 31- and 41-point Gauss-Hermite tables
(quadrature_tables/gaussHermite*.dat, kept here as tests/golden/ fixtures)
that none of its code reads; its 3-D rapidity integral uses the tan grid
(src/FreeStream.cpp:174-175, src/LandauMatch.cpp:93-95).  The library's rule:
  * nodes and weights (fsm_gauss_hermite) against the reference's tables to
    the digits they print, and against numpy.polynomial.hermite.hermgauss;
  * the oracle under the same rule (nodes from numpy) at DTAU = 0 against the
    closed-form rapidity moments of the density cosh^2(u) exp(-u^2/2 sigma^2):
    T^tt = e0, tau^2 T^nn = e0 tanh(sigma^2), T^xx + T^yy = 2 e0 / (1 +
    exp(2 sigma^2)), T^tn = 0, with 21 nodes where the tan grid needs 401 for
    2e-3; the check fails when the density keeps one cosh.
"""
import os

import numpy as np
import pytest

import fsm_amd
import fsm_amd.workloads as W
import oracle as O

HB = float(np.float32(0.197326938))
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _table(n):
    rows = [line.split() for line in open(os.path.join(GOLDEN, "gaussHermite%dpt.dat" % n)) if line.strip()]
    return [r[0] for r in rows], [r[1] for r in rows]


def _printed_tol(text):
    """One unit in the last printed digit of a table entry (the tables are not
    always correctly rounded in their last digit: -1.593885861 in the 31-point
    table is -1.59388586047...)."""
    t = text.lower()
    mant, _, exp = t.partition("e")
    dec = len(mant.split(".")[1]) if "." in mant else 0
    return 10.0 ** (-dec + (int(exp) if exp else 0))


@pytest.mark.parametrize("n", [21, 31, 41])
def test_nodes_and_weights_match_the_reference_tables(n):
    x, w = fsm_amd.gauss_hermite(n)
    tx, tw = _table(n)
    assert len(tx) == n
    for k in range(n):
        assert abs(x[k] - float(tx[k])) <= _printed_tol(tx[k]) + 1e-15, (n, k, x[k], tx[k])
        assert abs(w[k] - float(tw[k])) <= _printed_tol(tw[k]) + 1e-15, (n, k, w[k], tw[k])
    xr, wr = np.polynomial.hermite.hermgauss(n)
    assert np.abs(x - xr).max() < 1e-12
    assert np.abs(w / wr - 1.0).max() < 1e-11
    assert abs(w.sum() - np.sqrt(np.pi)) < 1e-13


@pytest.mark.parametrize("n", [1, 2, 3, 8, 64, 200])
def test_rule_integrates_polynomials_exactly(n):
    x, w = fsm_amd.gauss_hermite(n)
    from math import gamma, sqrt

    kmax = min(2 * n, 61)  # x^k stays finite for the largest nodes
    for k in range(0, kmax, max(1, kmax // 7)):
        exact = 0.0 if k % 2 else gamma((k + 1) / 2.0)
        got = float(np.sum(w * x ** k))
        # relative to the sum of magnitudes: odd moments cancel exactly
        scale = float(np.sum(w * np.abs(x) ** k))
        assert abs(got - exact) <= 1e-12 * scale, (n, k, got, exact, scale)


def test_rule_rejects_bad_sizes():
    with pytest.raises(fsm_amd.FsmError):
        fsm_amd.gauss_hermite(0)
    with pytest.raises(fsm_amd.FsmError):
        fsm_amd.gauss_hermite(201)


def _dtau0_moments(sigma, precision, nrap=21):
    p = W.base_params(DIM_X=9, DIM_Y=9, DIM_ETA=7, DIM_RAP=nrap, DIM_PHIP=8, DX=0.1, DY=0.1, DETA=0.1,
                      TAU0=0.5, DTAU=0.0, SIGMA=sigma)
    e = np.full(9 * 9 * 7, 2.0, dtype=np.float32)
    q = O.params_finalize(p, e)
    r = O.run(q, e, precision, rapidity="gauss_hermite")
    return r["T"].reshape(10, 9, 9, 7)[:, 4, 4, 3], 2.0 / HB + 1e-7, float(q.TAU)


@pytest.mark.parametrize("sigma", [0.5, 1.0, 1.4])
@pytest.mark.parametrize("precision", ["double", "float"])
def test_oracle_gauss_hermite_dtau_zero_moments(sigma, precision):
    T, e0, tau = _dtau0_moments(sigma, precision)
    # 21 nodes: 1e-15 at sigma = 1, 5e-9 at sigma = 1.4 in FP64; the float
    # path's own rounding is ~1e-6
    tol = 5e-8 if precision == "double" else 3e-6
    assert abs(T[0] / e0 - 1.0) < tol, T[0] / e0
    assert abs(tau * tau * T[9] / T[0] - np.tanh(sigma ** 2)) < tol
    assert abs((T[4] + T[7]) / T[0] - 2.0 / (1.0 + np.exp(2 * sigma ** 2))) < tol
    assert abs(T[3]) < tol * T[0] / tau and abs(T[1]) < tol * T[0]


def test_gauss_hermite_moment_pin_catches_a_single_cosh():
    with O.mutation(O.MUTATIONS["rap_factor_cosh"]):
        T, e0, tau = _dtau0_moments(1.0, "double")
    assert abs(T[0] / e0 - 1.0) > 1e-3


def test_oracle_rule_is_restored_after_use():
    p, e = W.small_3d(nx=15, neta=9, nrap=5, nphi=8)
    q = O.params_finalize(p, e)
    a = O.run(q, e, "double")["T"]
    O.run(q, e, "double", rapidity="gauss_hermite")
    b = O.run(q, e, "double")["T"]
    assert np.array_equal(a, b)
