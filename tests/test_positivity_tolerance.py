"""The mixed path's spline positivity rule against the reference's bilinear
fallback (src/FreeStream.cpp:245-246: a sample takes the bicubic value when it
is > 0, else bilinear).

The MFMA form cannot choose per sample, so the spline kernel flags a spline
tile when a knot's Bernstein control points go below -1e-5 of the largest
knot of the tile's band region (the float rounding noise of the knot slopes
is ~1e-6 of that value), and MFMA tiles reading a flagged knot run the FP32
sample loop.  A patch whose true minimum lies in (-1e-5 max, 0] is therefore
kept on the MFMA form, with bicubic values where the reference takes
bilinear.  The case below builds such patches on purpose: a narrow hotspot
(15 GeV/fm^3, width 0.3 fm) and, 2 fm away where the hotspot has decayed to
the ingest floor, a 3x3-cell step of height 1e-3 fm^-4 -- the spline
undershoots to about -2.6e-4 next to the step, inside the tolerance of the
band's largest knot (76 fm^-4).  The CPU test shows that the case is
exercised (negative control points in tiles the rule accepts, and samples
where the reference's bicubic value is <= 0); the GPU test bounds what the
MFMA form does there against the reference float path."""
import numpy as np
import pytest

import fsm_amd
import fsm_amd.workloads as W
import oracle as O
from parity import HBARC, TOL, output_errs, tmunu_err, tmunu_local_err, well_conditioned

POS_EPS = 1.0e-5  # the spline kernel's tolerance, relative to the band's largest knot
STEP = 1.0e-3     # fm^-4
STEP_AT = (2.0, 0.0)


def positivity_dip_event():
    n, d = 101, 0.1
    p = W.base_params(DIM_X=n, DIM_Y=n, DIM_PHIP=51, DX=d, DY=d, TAU0=0.5, DTAU=0.5)
    x = W.grid_coords(n, d).astype(np.float64)
    X, Y = np.meshgrid(x, x, indexing="ij")
    e = 15.0 * np.exp(-(X ** 2 + Y ** 2) / (2 * 0.3 ** 2))
    step = (np.abs(X - STEP_AT[0]) <= 0.1001) & (np.abs(Y - STEP_AT[1]) <= 0.1001)
    e = e + step * STEP * HBARC
    return p, e.astype(np.float32).ravel()


def _knots(q, e):
    """Knot values (the reference's ingest, src/FreestreamMilne.cpp:251-271)
    and the double spline's slopes in cell units."""
    nx, ny = q.DIM_X, q.DIM_Y
    x = W.grid_coords(nx, q.DX).astype(np.float64)
    y = W.grid_coords(ny, q.DY).astype(np.float64)
    z = (e.reshape(nx, ny) / np.float32(HBARC) + np.float32(1e-7)).astype(np.float32).astype(np.float64)
    zx, zy, zxy = O.spline_derivs(x, y, z)
    return x, y, z, zx * q.DX, zy * q.DY, zxy * q.DX * q.DY


def _control_point_min(z, zx, zy, zxy):
    a, b, c = zx / 3.0, zy / 3.0, zxy / 9.0
    return np.minimum.reduce([z, z + a, z - a, z + b, z - b, z + a + b + c, z - a - b + c, z + a - b - c,
                              z - a + b - c])


def _flagged(z, cp):
    """The spline kernel's flag per 16 x 32 knot tile: a knot whose control
    points go below -POS_EPS times the largest |z| of the tile's band region
    (rows 16 i - 16 .. 16 i + 31, columns 32 j - 16 .. 32 j + 47)."""
    nx, ny = z.shape
    fx, fy = (nx + 15) // 16, (ny + 31) // 32
    bad = np.zeros((fx, fy), bool)
    for i in range(fx):
        for j in range(fy):
            zm = np.abs(z[max(0, 16 * i - 16):16 * i + 32, max(0, 32 * j - 16):32 * j + 48]).max()
            bad[i, j] = (cp[16 * i:16 * i + 16, 32 * j:32 * j + 32] < -POS_EPS * zm).any()
    return bad


def test_dip_case_exercises_the_tolerance():
    p, e = positivity_dip_event()
    q = fsm_amd.params_finalize(p, e)
    x, y, z, zx, zy, zxy = _knots(q, e)
    cp = _control_point_min(z, zx, zy, zxy)
    bad = _flagged(z, cp)
    assert not bad.any()  # every tile stays on the MFMA form
    neg = cp < 0
    assert neg.sum() >= 50, neg.sum()  # ... although these knots' patches can go negative
    assert cp.min() > -POS_EPS * z.max() and cp.min() < -1e-4
    # samples of cells around the step whose bicubic value the reference finds <= 0
    ix0 = int(round((STEP_AT[0] - x[0]) / q.DX))
    iy0 = int(round((STEP_AT[1] - y[0]) / q.DY))
    dtau = float(q.TAU) - float(q.TAU0)
    phi = np.arange(q.DIM_PHIP) * (2 * np.pi / q.DIM_PHIP)
    cells = [(i, j) for i in range(ix0 - 8, ix0 + 9) for j in range(iy0 - 8, iy0 + 9)]
    px = np.array([x[i] - np.cos(f) * dtau for i, _ in cells for f in phi])
    py = np.array([y[j] - np.sin(f) * dtau for _, j in cells for f in phi])
    v = O.spline_eval(x, y, z, px, py)
    n_bilinear = int((v <= 0).sum())
    print("knots with negative control points %d (min %.3g, tolerance %.3g); samples the reference takes "
          "bilinear near the step: %d of %d" % (neg.sum(), cp.min(), -POS_EPS * z.max(), n_bilinear, v.size))
    assert n_bilinear > 0


@pytest.mark.gpu
def test_mixed_path_on_dip_case():
    """The MFMA form keeps the bicubic value on those samples; the result stays
    within the parity bar against the reference float path, and the local
    error next to the step is reported against its own T^{tau tau} scale."""
    import torch

    p, e = positivity_dip_event()
    q = fsm_amd.params_finalize(p, e)
    print("positivity dip case DIM", q.DIM, "DIM_PHIP", q.DIM_PHIP, flush=True)
    ctx = fsm_amd.Context(0)
    ctx.set_path(fsm_amd.PATH_MIXED)
    D = q.DIM
    d_e = torch.from_numpy(e).cuda()
    d_T = torch.empty(10 * D, dtype=torch.float32, device="cuda")
    d_out = torch.empty(17 * D, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream()
    ctx.stream_tmunu_device(q, d_e.data_ptr(), d_T.data_ptr(), s)
    ctx.run_device(q, d_e.data_ptr(), d_out.data_ptr(), s)
    torch.cuda.synchronize()
    T, out = d_T.cpu().numpy().reshape(10, D), d_out.cpu().numpy().reshape(17, D)
    ctx.close()
    ref = O.run(q, e, "float")
    T00max = np.abs(ref["T"][0]).max()
    tf = tmunu_err(T, ref["T"])
    tl = tmunu_local_err(T, ref["T"], ref["out"])
    errs = output_errs(out, ref["out"], T00max, well_conditioned(ref["T"], q.TAU, "mixed"))
    # cells whose samples reach the step (within the streaming distance + 0.3 fm of it), on the side away
    # from the hotspot, where the step's matter dominates the hotspot's tail
    x = W.grid_coords(q.DIM_X, q.DX).astype(np.float64)
    X, Y = np.meshgrid(x, x, indexing="ij")
    near = (((X - STEP_AT[0]) ** 2 + (Y - STEP_AT[1]) ** 2 <= (float(q.TAU) - float(q.TAU0) + 0.3) ** 2) &
            (X >= STEP_AT[0])).ravel()
    dT = np.abs(T.astype(np.float64) - ref["T"])[:, near].max()
    t00_near = np.abs(ref["T"][0, near]).max()
    print("T err %.3g (of max T^tt), per cell %.3g; around the step: |dT| %.3g = %.3g of max T^tt, %.3g of the "
          "region's own largest T^tt (%.3g)" % (tf, tl, dT, dT / T00max, dT / t00_near, t00_near), errs, flush=True)
    assert tf <= TOL and tl <= TOL
    assert dT <= TOL * T00max
    for k in ("e", "p", "u", "pi", "Pi"):
        assert errs[k] <= TOL, (k, errs[k])
