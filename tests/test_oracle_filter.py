"""The oracle's 6th-order numerical filter (diff_6th_opt, dynamics/
explicit_diffusion.py:159-223, applied at runtime/operational_mode.py:
2870-2876), pinned by closed forms.

The filter's flux difference dflux(i+1) - dflux(i) is the sixth difference
of the field, so on a sampled sinusoid exp(i theta n) it multiplies the field
by (2 cos theta - 2)^3 = -64 sin^6(theta / 2); each flux is the fourth
difference of the face gradients, 16 sin^4(theta / 2) times the gradient,
never up-gradient, so the monotonic limiter leaves a sinusoid alone.  No
reference vector covers the filter; these identities and a known answer of
the limiter pin it.  (The limiter switches fluxes off where a face gradient
changes sign, so with it on the step is not differentiable there and
central differences of the step do not converge; the gradient oracle is
checked with the filter off.)
"""

import numpy as np
import pytest

from oracle import oracle
from wrf_gpu_amd import ic

COEF = 0.12 * 0.015625 / (2.0 * 6.0)


@pytest.mark.parametrize("n,m", [(16, 1), (16, 3), (24, 5), (13, 2)])
def test_sinusoid_along_x_is_damped_by_the_closed_form(n, m):
    th = 2 * np.pi * m / n
    x = np.cos(th * np.arange(n) + 0.3)
    f = np.tile(x, (2, 1, 1))
    got = oracle.sixth_tendency(f, COEF)
    want = -64.0 * np.sin(th / 2) ** 6 * COEF * f
    assert np.allclose(got, want, rtol=0, atol=1e-13 * COEF * 64)


def test_sinusoid_product_in_x_and_y():
    ny, nx, my, mx = 12, 20, 2, 3
    ty, tx = 2 * np.pi * my / ny, 2 * np.pi * mx / nx
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    f = (np.sin(ty * j + 0.2) * np.sin(tx * i + 0.7))[None]
    got = oracle.sixth_tendency(f, COEF)
    want = -64.0 * (np.sin(tx / 2) ** 6 + np.sin(ty / 2) ** 6) * COEF * f
    assert np.allclose(got, want, rtol=0, atol=1e-13 * COEF * 128)


def test_single_row_has_no_y_part_and_lines_conserve():
    rng = np.random.default_rng(4)
    f = rng.standard_normal((3, 1, 17))
    t = oracle.sixth_tendency(f, COEF)
    # one row: x only; the flux differences telescope, so each periodic line sums to zero
    assert np.allclose(t.sum(axis=2), 0.0, atol=1e-12)
    f2 = rng.standard_normal((2, 9, 11))
    t2 = oracle.sixth_tendency(f2, COEF)
    assert abs(t2.sum()) < 1e-11


def test_limiter_on_a_step_reduces_to_a_second_order_exchange():
    """A periodic step (0 below n/2, 1 above) has face gradients +1 at face
    n/2 and -1 at face 0, zero elsewhere.  The raw fluxes are the fourth
    difference of those gradients: 6g at the jump faces and -4g, +g on the
    faces beside them, where the gradient is zero, so the limiter zeroes
    every flux but the two at the jumps.  What is left moves 6 coef from the
    high cell to the low cell at each jump."""

    n = 16
    x = np.where(np.arange(n) < n // 2, 0.0, 1.0)
    t = oracle.sixth_tendency(x[None, None], COEF)[0, 0]
    want = np.zeros(n)
    want[n // 2 - 1], want[n // 2] = 6 * COEF, -6 * COEF
    want[n - 1], want[0] = -6 * COEF, 6 * COEF
    assert np.array_equal(t, want)


def test_filter_changes_the_step_and_keeps_it_finite():
    case = ic.stratified_terrain_3d(nx=16, ny=12, nz=8, dx=3000.0, u_sigma=1.0, sigma=0.5)
    a = oracle.run(case.metrics, case.namelist, case.state, 2)
    b = oracle.run(case.metrics, case.namelist.replace(diff_6th_opt=2), case.state, 2)
    assert all(np.isfinite(v).all() for v in b.values())
    assert not np.array_equal(a["theta"], b["theta"])
