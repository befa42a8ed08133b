"""Quad coefficient records ("qcoef", NDRR_LAYOUT_QCOEF) as the tests emulate and decode them
(tests.helpers.quad_coef_records / qcoef_to_coef), pinned on the CPU: in exact arithmetic the
interpolant lerp_x(Q(x0), Q(x0 + 1)) of the quad records is the cell's trilinear interpolant, i.e.
its coefficients are oracle.raymarch.coef_records' (a closed form, not the same expression: the
quad form takes y and z differences first, the coefficient form x first), and evaluating it gives
the corner-form oracle's values and gradients."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import raymarch as orm
from tests.helpers import qcoef_to_coef, quad_coef_records


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_quad_coefficients_are_the_cells_trilinear_coefficients(seed):
    rng = np.random.default_rng(seed)
    vol = rng.standard_normal((5, 6, 7)).astype(np.float32)
    vol[2] = 0.0
    q = quad_coef_records(vol, np.float64)
    assert q.shape == (6, 7, 9, 4)
    got = qcoef_to_coef(q)
    want = orm.coef_records(vol, np.float64)
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_quad_coefficient_interpolant_matches_the_corner_oracle():
    """Values and voxel-space gradients at random points (inside, on faces, outside the volume)."""
    rng = np.random.default_rng(5)
    vol = rng.uniform(0, 2, (4, 5, 6))
    coef = qcoef_to_coef(quad_coef_records(vol, np.float64))
    D, H, W = vol.shape
    p = np.column_stack([rng.uniform(-1.5, W + 0.5, 3000), rng.uniform(-1.5, H + 0.5, 3000),
                         rng.uniform(-1.5, D + 0.5, 3000)])
    p[:300] = np.round(p[:300])  # faces and corners
    val, gx, gy, gz = orm._coef_eval(coef, p)
    vals, f, _ = orm._corners(vol, p)
    np.testing.assert_allclose(val, orm._lerp(vals, f), rtol=0, atol=1e-12)
    # gradient: the corner form's edge differences, lerped over the other two fractions
    fx, fy, fz = f[..., 0], f[..., 1], f[..., 2]
    dx = [[vals[a, b, 1] - vals[a, b, 0] for b in (0, 1)] for a in (0, 1)]
    want_gx = (dx[0][0] * (1 - fy) + dx[0][1] * fy) * (1 - fz) + (dx[1][0] * (1 - fy) + dx[1][1] * fy) * fz
    np.testing.assert_allclose(gx, want_gx, rtol=0, atol=1e-12)
    inside = (p >= 0).all(axis=1) & (p[:, 0] <= W - 1) & (p[:, 1] <= H - 1) & (p[:, 2] <= D - 1)
    assert inside.sum() > 500
