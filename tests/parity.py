"""Parity metrics shared by the GPU tests.

Tolerances (stated by the task for this path):
  * T^{mu nu}: |dT| <= 1e-4 * max T^{tau tau} (fm^-4), every cell and component;
  * e: |de| <= 1e-4 * e_ref and u^mu: |du^mu| <= 1e-4 * u^tau_ref on cells with
    e_ref > 1e-3 * max e_ref whose Landau matching is well conditioned (below);
  * pi^{mu nu}, Pi (GeV/fm^3): |d| <= 1e-4 * max T^{tau tau} * hbarc on the
    same cells.

Conditioning.  Where T^{mu nu} is close to a null tensor (a few streaming
directions dominate, at grid edges and in dilute cells) the Landau eigenvalue
is a small difference of large numbers: a change of 1e-6 T^{tau tau} in the
components -- the rounding level of the reference's own float sums over
N_phi*N_rap positive samples, which is set by T^{tau tau} and not by each
component's own size -- can move e by 1e-2 or more.  The reference float
path does not determine e or u there to 1e-4, so the e/u/pi comparison uses
only cells with kappa = max |de/e| / (|dT| / T^{tau tau}) <= KAPPA_MAX[path] (kappa
from oracle.landau_condition: first-order worst case over perturbations of
every component by up to 1e-6 T^{tau tau}, FP64).  The stage-B tests compare the GPU Landau kernels on
identical input T with no such exclusion.

The excluded cells are not left unchecked: `excluded_cells_check` compares
GPU and reference float path on them against the FP64 oracle, and the GPU's
largest e and u errors there must be no more than twice the reference float
path's own (plus the tolerance).  Their number is bounded per case
(`EXCLUDED_BUDGET`): in 2-D they are the dilute cells of the outer ring (about
1% of the cells with e > 1e-3 e_max); in 3-D mostly the eta-edge planes, where
the reference's eta mask drops every sample streaming in from outside the
grid, leaving a one-sided, nearly null tensor (up to a fifth of the cells of a
21-plane grid, a few percent of a 101-plane one).
"""
import numpy as np

HBARC = float(np.float32(0.197326938))
TOL = 1.0e-4
# Conditioning threshold per quadrature path.  Two tensors that are each off
# by delta (relative to T^{tau tau}) can give e values up to kappa*(2 delta)
# apart.  The reference float path's own rounding is ~1e-6; the GPU FP32 path
# is ~3e-7 off the FP64 oracle, the mixed path (fp16 hi+lo residual) ~1e-6, so
# the 1e-4 bar on e and u is met by construction for kappa up to ~75 and ~50.
KAPPA_MAX = {"fp32": 100.0, "mixed": 50.0}


def tmunu_err(T_gpu, T_ref, keep=None):
    T_gpu = np.asarray(T_gpu, dtype=np.float64).reshape(10, -1)
    T_ref = np.asarray(T_ref, dtype=np.float64).reshape(10, -1)
    scale = np.abs(T_ref[0]).max()
    d = np.abs(T_gpu - T_ref)
    if keep is not None:
        d = d[:, keep]
    return d.max() / scale


def tmunu_local_err(T_gpu, T_ref, out_ref):
    """Largest |dT^{mu nu}| relative to the cell's own T^{tau tau}, over the
    cells with e > 1e-3 e_max (out_ref: the reference's 17 outputs).  The
    global scale of tmunu_err hides an error in a dilute cell; this one does
    not."""
    T_gpu = np.asarray(T_gpu, dtype=np.float64).reshape(10, -1)
    T_ref = np.asarray(T_ref, dtype=np.float64).reshape(10, -1)
    e = np.asarray(out_ref, dtype=np.float64).reshape(17, -1)[0]
    m = e > 1e-3 * e.max()
    d = np.abs(T_gpu - T_ref)[:, m].max(axis=0)
    t00 = np.abs(T_ref[0, m])
    # a cell that keeps no sample (T = 0, the reference's fallback frame) must be exactly 0
    rel = np.where(t00 > 0, d / np.where(t00 > 0, t00, 1.0), np.where(d > 0, np.inf, 0.0))
    return float(rel.max()) if rel.size else 0.0


def interior_mask(q, margin: int = 2):
    """Cells at least `margin` cells from every grid edge (x, y and, in 3-D,
    eta).  Next to an edge a sample whose exact position is a grid line sits on
    the reference's mask boundary (floor index in [1, N-2]); float and double
    rounding then keep or drop it differently."""
    m = np.zeros((q.DIM_X, q.DIM_Y, q.DIM_ETA), dtype=bool)
    if q.DIM_ETA > 1:
        m[margin:q.DIM_X - margin, margin:q.DIM_Y - margin, margin:q.DIM_ETA - margin] = True
    else:
        m[margin:q.DIM_X - margin, margin:q.DIM_Y - margin, :] = True
    return m.ravel()


def ring_interior_mask(q, margin: int = 2):
    """2-D: cells whose every streaming sample lies `margin` or more cells
    inside the grid (the ring of radius DTAU/DX around the cell).  Float and
    FP64 keep different samples only at the mask boundary, so this is where
    the FP64 comparison is meaningful; with DTAU/DX a whole number the
    boundary samples reach exactly DTAU/DX cells in from the edges."""
    r = int(np.ceil(float(q.DTAU) / min(float(q.DX), float(q.DY))))
    return interior_mask(q, margin + r)


def condition_numbers(T_ref, tau):
    import oracle as O

    return O.landau_condition(np.asarray(T_ref).reshape(10, -1).T, np.float32(tau))


def well_conditioned(T_ref, tau, path="fp32", kappa=None):
    """Cells whose Landau matching the reference float path determines to
    the tolerance on `path` (kappa <= KAPPA_MAX[path]); `kappa` = precomputed
    condition_numbers."""
    ke, ku = condition_numbers(T_ref, tau) if kappa is None else kappa
    k = KAPPA_MAX[path]
    return (ke <= k) & (ku <= k)


def output_errs(out_gpu, out_ref, T00max, cond=None):
    """out_*: (17, DIM) in output_to_vectors order.  Returns dict of errors in
    units of the tolerance scales above."""
    g = np.asarray(out_gpu, dtype=np.float64).reshape(17, -1)
    r = np.asarray(out_ref, dtype=np.float64).reshape(17, -1)
    e_ref = r[0]
    mask = e_ref > 1e-3 * e_ref.max()
    if cond is not None:
        mask = mask & cond
    errs = {}
    errs["e"] = (np.abs(g[0] - r[0])[mask] / e_ref[mask]).max()
    # p relative to itself, or to e where the EoS gives p = 0 (EOS_TYPE 3)
    p_ref = np.abs(r[1][mask])
    errs["p"] = (np.abs(g[1] - r[1])[mask] / np.where(p_ref > 0, p_ref, e_ref[mask])).max()
    ut = r[2][mask]
    errs["u"] = max((np.abs(g[2 + k] - r[2 + k])[mask] / ut).max() for k in range(4))
    scale = T00max * HBARC
    errs["pi"] = np.abs(g[6:16, mask] - r[6:16, mask]).max() / scale
    errs["Pi"] = np.abs(g[16, mask] - r[16, mask]).max() / scale
    errs["cells"] = int(mask.sum())
    errs["excluded_ill_conditioned"] = int(((e_ref > 1e-3 * e_ref.max()) & ~mask).sum())
    return errs


# Largest fraction of the cells with e > 1e-3 e_max that the conditioning
# filter may leave out of the strict e/u comparison, per path (2-D; 3-D grids
# of 21 or fewer eta planes; other 3-D grids).  The fraction grows with the
# streaming ring's radius DTAU/DX: over the 512-event 2-D mix the largest is
# 11% at kappa 100 and 19% at 50 (301^2 events with DTAU/DX ~ 17), C1 (16.7)
# has 1.1% and 3.8%; C2-style 3-D events 4% and 9%.
EXCLUDED_BUDGET = {"fp32": {"2d": 0.15, "3d_thin": 0.25, "3d": 0.12},
                   "mixed": {"2d": 0.25, "3d_thin": 0.35, "3d": 0.20}}


def excluded_budget(q, path="fp32"):
    b = EXCLUDED_BUDGET[path]
    if q.DIM_ETA == 1:
        return b["2d"]
    return b["3d_thin"] if q.DIM_ETA <= 21 else b["3d"]


def fallback_cells(out):
    """Cells whose outputs are the reference's fallback frame."""
    o = np.asarray(out, dtype=np.float64).reshape(17, -1)
    e_fb = float(np.float32(1.0e-3)) * HBARC
    return (o[0] == e_fb) & (o[2] == 1.0) & (o[3] == 0.0) & (o[4] == 0.0) & (o[5] == 0.0)


def excluded_cells_check(out_gpu, out_float, out_double, cond, q, path="fp32", T_float=None, T_double=None,
                         budget=None):
    """On the cells with e > 1e-3 e_max that `cond` leaves out: largest
    relative e error and u^mu error (in units of u^tau) of the GPU and of the
    reference float path, both against the FP64 oracle.  Asserts the count
    budget and that the GPU is within twice the float path's error plus the
    tolerance; returns the figures for the test log.  `budget` replaces the
    per-path budget for a case built to be one-sided (a streaming ring wider
    than the grid).

    With T_float and T_double the comparison uses only the cells whose tensor
    is the same on both oracle paths (every component within 1e-5 of the
    cell's T^{tau tau}).  Where a sample sits on a mask boundary (a whole
    number of cells from the grid edge) float and FP64 keep different samples
    and build different tensors; e and u of two different, nearly null tensors
    say nothing about either implementation.  Those cells are counted."""
    g = np.asarray(out_gpu, dtype=np.float64).reshape(17, -1)
    f = np.asarray(out_float, dtype=np.float64).reshape(17, -1)
    d = np.asarray(out_double, dtype=np.float64).reshape(17, -1)
    big = f[0] > 1e-3 * f[0].max()
    exc = big & ~cond
    res = {"excluded": int(exc.sum()), "frac": float(exc.sum()) / max(int(big.sum()), 1)}
    assert res["frac"] <= (excluded_budget(q, path) if budget is None else budget), res
    if T_float is not None and T_double is not None:
        tf = np.asarray(T_float, dtype=np.float64).reshape(10, -1)
        td = np.asarray(T_double, dtype=np.float64).reshape(10, -1)
        same = (np.abs(tf - td).max(axis=0) <= 1e-5 * np.abs(td[0]))
        res["tensor_differs"] = int((exc & ~same).sum())
        exc = exc & same
    # Where either oracle path takes the reference's fallback (e = 1e-3 fm^-4,
    # u = (1, 0, 0, 0); src/LandauMatch.cpp:283-291) the reported e is not an
    # eigenvalue of T: on these nearly null, single-beam cells (kappa ~ 1e6)
    # the accept/fallback decision flips with the 7th digit of T (the
    # oracle's own FP64 solve of the GPU's tensor finds the GPU's frame).  In
    # a dilute event the fallback value can exceed 1e-3 e_max, which is how
    # such cells pass the e filter.  They are counted, not compared.
    fb = fallback_cells(f) | fallback_cells(d)
    res["fallback_decision"] = int((exc & fb).sum())
    exc = exc & ~fb
    if not exc.any():
        return res
    ed, ut = d[0][exc], d[2][exc]
    for name, a in (("gpu", g), ("float", f)):
        res[name + "_e"] = float((np.abs(a[0][exc] - ed) / ed).max())
        res[name + "_u"] = float(max((np.abs(a[2 + k][exc] - d[2 + k][exc]) / ut).max() for k in range(4)))
    assert res["gpu_e"] <= 2.0 * res["float_e"] + TOL, res
    assert res["gpu_u"] <= 2.0 * res["float_u"] + TOL, res
    return res
