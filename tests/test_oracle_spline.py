"""Pins of the oracle's natural bicubic spline (restatement of GSL's
gsl_interp2d_bicubic as called at src/FreeStream.cpp:121-153, :240).

Independent references: scipy's natural cubic splines applied along each axis
(the tensor-product natural spline is unique), affine reproduction, and the
interpolation property."""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

import oracle as O


def _grid(rng, nx, ny, uniform):
    if uniform:
        return np.arange(nx) * 0.1 - 0.3, np.arange(ny) * 0.07 + 0.2
    return np.sort(rng.uniform(0, 5, nx)), np.sort(rng.uniform(-1, 3, ny))


@pytest.mark.parametrize("uniform", [True, False])
def test_knot_derivatives_match_scipy_natural(uniform):
    rng = np.random.default_rng(1)
    x, y = _grid(rng, 9, 13, uniform)
    z = rng.normal(size=(9, 13))
    zx, zy, zxy = O.spline_derivs(x, y, z)
    ref_zx = np.stack([CubicSpline(x, z[:, j], bc_type="natural")(x, 1) for j in range(13)], axis=1)
    ref_zy = np.stack([CubicSpline(y, z[i], bc_type="natural")(y, 1) for i in range(9)], axis=0)
    ref_zxy = np.stack([CubicSpline(y, ref_zx[i], bc_type="natural")(y, 1) for i in range(9)], axis=0)
    s = np.abs(z).max()
    assert np.abs(zx - ref_zx).max() < 1e-10 * s * 100
    assert np.abs(zy - ref_zy).max() < 1e-10 * s * 100
    assert np.abs(zxy - ref_zxy).max() < 1e-9 * s * 100


def test_eval_matches_tensor_product_natural_spline():
    rng = np.random.default_rng(2)
    x, y = _grid(rng, 8, 10, False)
    z = rng.normal(size=(8, 10))
    px = rng.uniform(x[0], x[-1], 300)
    py = rng.uniform(y[0], y[-1], 300)
    v = O.spline_eval(x, y, z, px, py)
    ref = []
    for a, b in zip(px, py):
        col = np.array([CubicSpline(y, z[i], bc_type="natural")(b) for i in range(len(x))])
        ref.append(CubicSpline(x, col, bc_type="natural")(a))
    assert np.abs(v - np.array(ref)).max() < 1e-12


def test_affine_data_reproduced_exactly():
    x = np.arange(12) * 0.06 - 0.33
    y = np.arange(9) * 0.06 + 0.1
    z = 2.0 + 3.0 * x[:, None] - 1.5 * y[None, :] + 0.7 * x[:, None] * y[None, :]
    rng = np.random.default_rng(3)
    px = rng.uniform(x[0], x[-1], 100)
    py = rng.uniform(y[0], y[-1], 100)
    v = O.spline_eval(x, y, z, px, py)
    assert np.abs(v - (2.0 + 3.0 * px - 1.5 * py + 0.7 * px * py)).max() < 1e-12


def test_interpolates_knots_and_natural_end_curvature():
    rng = np.random.default_rng(4)
    x = np.arange(7) * 1.0
    y = np.arange(6) * 0.5
    z = rng.normal(size=(7, 6))
    px, py = np.meshgrid(x[:-1], y[:-1], indexing="ij")
    v = O.spline_eval(x, y, z, px.ravel(), py.ravel())
    assert np.abs(v - z[:-1, :-1].ravel()).max() < 1e-12
    # natural end: second x-derivative vanishes at x[0] along a knot row
    h = 1e-4
    f = O.spline_eval(x, y, z, np.array([x[0], x[0] + h, x[0] + 2 * h]), np.full(3, y[2]))
    assert abs((f[0] - 2 * f[1] + f[2]) / h ** 2) < 1e-2 * np.abs(z).max()
