"""Integration grids without a GPU: the Lebedev rules integrate every
monomial up to their degree exactly (against the closed-form sphere moments),
and the Pople radial grid integrates known radial functions."""
import math

import numpy as np
import pytest

from pgxc import grid, lebedev


@pytest.mark.parametrize("n", [302, 194])
def test_lebedev_exact_to_degree(n):
    P, w = lebedev.rule(n)
    assert P.shape == (n, 3) and w.shape == (n,)
    assert np.allclose(np.linalg.norm(P, axis=1), 1.0, atol=1e-15)
    assert abs(w.sum() - 1.0) <= 1e-14 and np.all(w > 0)
    deg = lebedev.DEGREE[n]
    worst = 0.0
    for d in range(deg + 1):
        for a in range(d + 1):
            for b in range(d + 1 - a):
                c = d - a - b
                exact = lebedev.sphere_moment(a, b, c) if a % 2 == b % 2 == c % 2 == 0 else 0.0
                got = np.sum(w * P[:, 0] ** a * P[:, 1] ** b * P[:, 2] ** c)
                worst = max(worst, abs(got - exact))
    assert worst <= 5e-15
    # and the next even degree is not exact (the rule is no better than stated)
    d = deg + 1
    errs = [abs(np.sum(w * P[:, 0] ** a * P[:, 1] ** (d - a)) - lebedev.sphere_moment(a, d - a, 0))
            for a in range(0, d + 1, 2)]
    assert max(errs) > 1e-13  # far above the 5e-15 rounding floor


def test_sphere_moment_closed_form():
    # <x^2> = 1/3, <x^4> = 1/5, <x^2 y^2> = 1/15 over the unit sphere
    assert abs(lebedev.sphere_moment(2, 0, 0) - 1 / 3) < 1e-15
    assert abs(lebedev.sphere_moment(4, 0, 0) - 1 / 5) < 1e-15
    assert abs(lebedev.sphere_moment(2, 2, 0) - 1 / 15) < 1e-15


def test_radial_grid_integrates_gaussians():
    # int_0^inf r^2 exp(-a r^2) dr = sqrt(pi) / (4 a^1.5)
    r, w = grid.radial_grid(75, 1.134)
    for a in (0.3, 1.0, 10.0):
        exact = math.sqrt(math.pi) / (4 * a ** 1.5)
        assert abs(np.sum(w * np.exp(-a * r * r)) - exact) <= 1e-7 * exact
    assert r.max() < grid.RAD_LIMIT


def test_atomic_grid_sizes():
    for nang in (194, 302):
        p, w = grid.atomic_grid((0.0, 0.0, 0.0), 50, nang, 1.0)
        r, _ = grid.radial_grid(50, 1.0)
        assert p.shape == (len(r) * nang, 3)
        # the unit function integrates to the volume inside the last shell scale
        g = np.exp(-p[:, 0] ** 2 - p[:, 1] ** 2 - p[:, 2] ** 2)
        assert abs(np.sum(w * g) - math.pi ** 1.5) <= 1e-6 * math.pi ** 1.5
