"""ctypes loader of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product library.  The oracle is a
C++ restatement of the reference algorithm (see fsm_oracle.cpp's header for
the reference files it follows and for how each part is pinned).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

_LIB = None


def build(force: bool = False):
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-C", _HERE, "-s"])


def _lib():
    global _LIB
    if _LIB is None:
        build()
        lib = ctypes.CDLL(LIB_PATH)
        from fsm_amd import Params  # same struct layout as include/fsm.h

        P = ctypes.POINTER(Params)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        lib.oracle_params_default.argtypes = [P]
        lib.oracle_params_default.restype = None
        lib.oracle_params_finalize.argtypes = [P, fp, ctypes.c_long]
        lib.oracle_run.argtypes = [P, fp, fp, ctypes.c_int, dp, dp, dp, dp]
        lib.oracle_spline_derivs.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, dp, dp, dp]
        lib.oracle_spline_derivs.restype = None
        lib.oracle_spline_eval.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, ctypes.c_int, dp, dp, dp]
        lib.oracle_spline_eval.restype = None
        lib.oracle_eig4.argtypes = [dp, dp, dp, dp]
        lib.oracle_landau.argtypes = [ctypes.c_long, dp, ctypes.c_float, dp, dp]
        lib.oracle_landau.restype = None
        lib.oracle_num_threads.restype = ctypes.c_int
        _LIB = lib
    return _LIB


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def params_default():
    from fsm_amd import Params

    p = Params()
    _lib().oracle_params_default(ctypes.byref(p))
    return p


def params_finalize(p, e=None):
    q = p.copy()
    if e is None:
        st = _lib().oracle_params_finalize(ctypes.byref(q), None, 0)
    else:
        e = np.ascontiguousarray(e, dtype=np.float32)
        st = _lib().oracle_params_finalize(ctypes.byref(q), _f(e), e.size)
    if st != 0:
        raise ValueError("oracle_params_finalize status %d" % st)
    return q


class gauss_hermite:
    """Context manager switching the oracle's 3-D rapidity rule to the
    n-point Gauss-Hermite rule of the library's FSM_RAPIDITY_GAUSS_HERMITE
    extension (nodes eta + sqrt(2) SIGMA x_k, weights sqrt(2) SIGMA W_k).  The
    nodes and weights come from numpy.polynomial.hermite.hermgauss, not from
    the library.  The reference's tan grid is restored on exit."""

    def __init__(self, n: int):
        self.x, self.w = np.polynomial.hermite.hermgauss(int(n))

    def __enter__(self):
        lib = _lib()
        dp = ctypes.POINTER(ctypes.c_double)
        lib.oracle_set_rapidity_rule.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        x = np.ascontiguousarray(self.x, dtype=np.float64)
        w = np.ascontiguousarray(self.w, dtype=np.float64)
        if lib.oracle_set_rapidity_rule(1, x.size, _d(x), _d(w)) != 0:
            raise ValueError("oracle_set_rapidity_rule failed")
        return self

    def __exit__(self, *exc):
        _lib().oracle_set_rapidity_rule(0, 0, None, None)
        return False


def run(p, e, precision: str = "float", nB=None, want_T: bool = True, want_extra: bool = True,
        rapidity: str = "tan"):
    """Whole step.  precision 'float' = reference float path, 'double' = FP64 oracle.

    p must be finalized.  Returns dict with 'out' (17 x DIM, output_to_vectors
    order), 'T' (10 x DIM, fm^-4), 'e', 'u' (fm^-4 / dimensionless), optional
    'nB', 'V' (baryon), and 'stage_seconds'.  rapidity 'gauss_hermite' runs a
    3-D event under the Gauss-Hermite rule (see gauss_hermite)."""
    if rapidity == "gauss_hermite":
        with gauss_hermite(p.DIM_RAP):
            return run(p, e, precision, nB, want_T, want_extra)
    if rapidity != "tan":
        raise ValueError("rapidity rule %r" % rapidity)
    e = np.ascontiguousarray(e, dtype=np.float32).ravel()
    D = p.DIM
    if e.size != D:
        raise ValueError("size mismatch")
    out = np.empty((17, D))
    T = np.empty((10, D)) if want_T else None
    extra = np.zeros((10, D)) if want_extra else None
    secs = np.zeros(4)
    nb = None if nB is None else np.ascontiguousarray(nB, dtype=np.float32).ravel()
    st = _lib().oracle_run(ctypes.byref(p), _f(e), None if nb is None else _f(nb),
                           0 if precision == "float" else 1, _d(out),
                           None if T is None else _d(T), None if extra is None else _d(extra), _d(secs))
    if st != 0:
        raise ValueError("oracle_run status %d" % st)
    r = {"out": out, "T": T, "stage_seconds": secs}
    if extra is not None:
        r["e"] = extra[0]
        r["u"] = extra[1:5]
        if nb is not None:
            r["nB"] = extra[5]
            r["V"] = extra[6:10]
    return r


def spline_derivs(x, y, z):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    nx, ny = x.size, y.size
    zx, zy, zxy = (np.empty((nx, ny)) for _ in range(3))
    _lib().oracle_spline_derivs(nx, ny, _d(x), _d(y), _d(z), _d(zx), _d(zy), _d(zxy))
    return zx, zy, zxy


def spline_eval(x, y, z, px, py):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    px = np.ascontiguousarray(px, dtype=np.float64)
    py = np.ascontiguousarray(py, dtype=np.float64)
    out = np.empty(px.size)
    _lib().oracle_spline_eval(x.size, y.size, _d(x), _d(y), _d(z), px.size, _d(px), _d(py), _d(out))
    return out


def eig4(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    wr, wi, V = np.empty(4), np.empty(4), np.empty((4, 4))
    ok = _lib().oracle_eig4(_d(A), _d(wr), _d(wi), _d(V))
    return wr, wi, V, ok


def landau(T, tau: float):
    """T: (n, 10) doubles -> e (n,), u (n, 4) by the reference's selection rules."""
    T = np.ascontiguousarray(T, dtype=np.float64)
    n = T.shape[0]
    e, u = np.empty(n), np.empty((n, 4))
    _lib().oracle_landau(n, _d(T), tau, _d(e), _d(u))
    return e, u


def num_threads() -> int:
    return _lib().oracle_num_threads()


def pressure(e, eos: int, precision: str = "float"):
    """calculatePressure on an array of energy densities (fm^-4)."""
    lib = _lib()
    if not hasattr(lib, "_eos_ready"):
        lib.oracle_pressure.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double)]
        lib.oracle_pressure.restype = None
        lib._eos_ready = True
    e = np.ascontiguousarray(e, dtype=np.float64)
    out = np.empty_like(e)
    lib.oracle_pressure(eos, 0 if precision == "float" else 1, e.size, _d(e), _d(out))
    return out


def viscous_float(p, T, e, u, pr):
    """The reference float path's calculateBulkPressure and
    calculateShearViscTensor on given float planes: (Pi, pi[10])."""
    lib = _lib()
    from fsm_amd import Params

    fp = ctypes.POINTER(ctypes.c_float)
    lib.oracle_viscous_float.argtypes = [ctypes.POINTER(Params), fp, fp, fp, fp, fp, fp]
    lib.oracle_viscous_float.restype = None
    D = p.DIM
    a = [np.ascontiguousarray(x, dtype=np.float32).ravel() for x in (T, e, u, pr)]
    Pi = np.zeros(D, dtype=np.float32)
    pi = np.zeros(10 * D, dtype=np.float32)
    lib.oracle_viscous_float(ctypes.byref(p), *[_f(x) for x in a], _f(Pi), _f(pi))
    return Pi, pi.reshape(10, D)


class mutation:
    """Context manager planting one mistake in the oracle (see the Mutation
    ids in fsm_oracle.cpp), used only by the tests that show a pin fails on
    it; the oracle is restored on exit."""

    def __init__(self, m: int):
        self.m = m

    def __enter__(self):
        lib = _lib()
        lib.oracle_set_mutation.argtypes = [ctypes.c_int]
        lib.oracle_set_mutation.restype = ctypes.c_int
        self.old = lib.oracle_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        _lib().oracle_set_mutation(self.old)
        return False


MUTATIONS = {
    "trilinear_corner": 1, "eta_new_sign": 2, "shift_no_tau0": 3, "jacobian_one_cos": 4, "3d_over_tau": 5,
    "charge_width": 6, "charge_norm": 7, "edep_no_hbarc": 8, "eos2_swap": 9, "hypertrig_tau": 10,
    "eta_mask_low": 11, "rap_factor_cosh": 12, "baryon_no_drap": 13,
}


def landau_condition(T, tau: float, ntrial: int = 4, rel: float = 1e-6, seed: int = 12345, normwise: bool = True):
    """Per-cell condition numbers (kappa_e, kappa_u) of the Landau matching:
    relative change of e (and of u^mu in units of u^tau) per unit relative
    perturbation of T.  T: (n, 10).
    normwise (default): first-order worst case over every component moving by
    up to rel * T^{tau tau} (1/tau per eta index), the scale of a quadrature
    sum's rounding error (one-sided differences per component); otherwise the
    largest response to `ntrial` random perturbations of each component
    relative to itself."""
    lib = _lib()
    if not hasattr(lib, "_cond_ready"):
        lib.oracle_landau_condition.argtypes = [ctypes.c_long, ctypes.POINTER(ctypes.c_double), ctypes.c_float,
                                                ctypes.c_int, ctypes.c_double, ctypes.c_uint, ctypes.c_int,
                                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        lib.oracle_landau_condition.restype = None
        lib._cond_ready = True
    T = np.ascontiguousarray(T, dtype=np.float64)
    n = T.shape[0]
    ke, ku = np.empty(n), np.empty(n)
    lib.oracle_landau_condition(n, _d(T), tau, ntrial, rel, seed, 1 if normwise else 0, _d(ke), _d(ku))
    return ke, ku


def _diag_lib():
    lib = _lib()
    if not hasattr(lib, "_diag_ready"):
        from fsm_amd import Params

        fp = ctypes.POINTER(ctypes.c_float)
        lib.oracle_diagnostics.argtypes = [ctypes.POINTER(Params), fp, fp, fp, fp, fp, fp]
        lib.oracle_diagnostics.restype = None
        lib.oracle_energy_sums.argtypes = [ctypes.POINTER(Params), fp, fp, fp, fp]
        lib.oracle_energy_sums.restype = None
        lib._diag_ready = True
    return lib


def diagnostics(p, e, u, pr, Pi, pi):
    """The reference's diagnostics on float hydro fields (see
    fsm_oracle.cpp): dict of R_Pi_Inv, R_pimunu_Inv, pi_dot_u (4, D),
    pi_mu_mu, beta (4, D), omega (6, D)."""
    D = p.DIM_X * p.DIM_Y * p.DIM_ETA
    f = [np.ascontiguousarray(a, dtype=np.float32).ravel() for a in (e, u, pr, Pi, pi)]
    out = np.zeros(17 * D, dtype=np.float32)
    _diag_lib().oracle_diagnostics(ctypes.byref(p), *[_f(a) for a in f], _f(out))
    out = out.reshape(17, D)
    return {"R_Pi_Inv": out[0], "R_pimunu_Inv": out[1], "pi_dot_u": out[2:6], "pi_mu_mu": out[6],
            "beta": out[7:11], "omega": out[11:17]}


def energy_sums(p, T00, e):
    """Serial float sums of the energy after streaming and inside the
    freeze-out surface (T00, e in fm^-4)."""
    T00 = np.ascontiguousarray(T00, dtype=np.float32)
    e = np.ascontiguousarray(e, dtype=np.float32)
    a, b = np.zeros(1, np.float32), np.zeros(1, np.float32)
    _diag_lib().oracle_energy_sums(ctypes.byref(p), _f(T00), _f(e), _f(a), _f(b))
    return float(a[0]), float(b[0])
