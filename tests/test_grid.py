"""Integration grids without a GPU: the Lebedev rules integrate every
monomial up to their degree exactly (against the closed-form sphere moments),
and the Pople radial grid integrates known radial functions."""
import math

import numpy as np
import pytest

from pgxc import grid, lebedev


@pytest.mark.parametrize("n", [302, 194])
def test_lebedev_exact_to_degree(n):
    P, w = lebedev.rule(n, "solved")
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


# ------------------------------------------------ the reference's grid, pinned
import itertools  # noqa: E402

import oracle  # noqa: E402
from pgxc import systems  # noqa: E402


def _octahedral_images(v):
    return {tuple(np.round([s * c for s, c in zip(sg, perm)], 12))
            for perm in itertools.permutations(v) for sg in itertools.product((1, -1), repeat=3)}


@pytest.mark.parametrize("n", [302, 194])
def test_tabulated_lebedev_rule(n):
    """The reference's 12-digit rule: unit points in closed octahedral
    orbits, exact on monomials to its degree to the table's digits, and the
    exactly solved rule of the same structure to those digits."""
    P, w = lebedev.rule(n)
    assert P.shape == (n, 3) and len({tuple(np.round(p, 9)) for p in P}) == n
    assert np.abs(np.linalg.norm(P, axis=1) - 1.0).max() <= 3e-16
    keys = {tuple(np.round(p, 12)) for p in P}
    for p, wt in zip(P, w):  # every image of a point is in the rule with its weight
        for q in _octahedral_images(p):
            assert q in keys
    deg, worst = lebedev.DEGREE[n], 0.0
    for d in range(deg + 1):
        for a in range(d + 1):
            for b in range(d + 1 - a):
                c = d - a - b
                exact = lebedev.sphere_moment(a, b, c) if a % 2 == b % 2 == c % 2 == 0 else 0.0
                worst = max(worst, abs(np.sum(w * P[:, 0] ** a * P[:, 1] ** b * P[:, 2] ** c) - exact))
    assert worst <= 1e-12
    Ps, ws = lebedev.rule(n, "solved")
    ka, kb = np.lexsort(np.round(P, 9).T), np.lexsort(np.round(Ps, 9).T)
    assert np.abs(P[ka] - Ps[kb]).max() <= 5e-12
    # 12 digits, except the 302 rule's a1 weight (8.54591172878e-4 against
    # 8.5459117238e-4 exactly: 5.9e-10 relative)
    assert np.abs(w[ka] / ws[kb] - 1).max() <= (1e-9 if n == 302 else 1e-11)


@pytest.mark.parametrize("n", [302, 194])
def test_oracle_lebedev_equals_product_rule(n):
    # two restatements of INIGRD's orbit expansion, point for point
    P, w = oracle.lebedev(n)
    Pp, wp = lebedev.rule(n)
    assert np.array_equal(P, Pp) and np.array_equal(w, wp)


def test_oracle_radial_grid_is_the_pople_map():
    """w_i = r(i)^2 dr/di of the continuous map r(i) = R (i/(n+1-i))^2,
    shells up to the last with r < 40 bohr, Gaussian moments."""
    for R, n in ((1.0015551316604694, 75), (1.1338359981061916, 85), (1.3228086644572234, 99)):
        r, w = oracle.radial_grid(n, R)
        m = len(r)
        assert r[-1] < 40.0 and (m == n or R * ((m + 1) / (n - m)) ** 2 >= 40.0)
        i = np.arange(1, m + 1, dtype=float)
        rr = lambda t: R * (t / (n + 1 - t)) ** 2  # noqa: E731
        h = 1e-5
        drdi = (rr(i + h) - rr(i - h)) / (2 * h)
        assert np.allclose(r, rr(i), rtol=1e-14, atol=0)
        assert np.abs(w / (r ** 2 * drdi) - 1).max() <= 1e-7
        for a in (0.5, 2.0):  # int r^2 exp(-a r^2) dr
            exact = math.sqrt(math.pi) / (4 * a ** 1.5)
            assert abs(np.sum(w * np.exp(-a * r * r)) - exact) <= 1e-6 * exact


def test_oracle_becke_partition():
    els, xyz = systems.water_cluster(2, seed=1)
    rng = np.random.default_rng(3)
    pts = xyz.mean(0) + rng.normal(scale=2.0, size=(200, 3))
    W = np.array([oracle.becke0(xyz, a, pts) for a in range(len(els))])
    assert np.abs(W.sum(0) - 1).max() <= 1e-14 and W.min() >= 0.0
    for a in range(len(els)):  # the host nucleus belongs to its own cell only
        assert oracle.becke0(xyz, a, xyz[a:a + 1])[0] == 1.0
    # brute force: product over every ordered pair of Becke's cut-off
    d = np.linalg.norm(xyz[:, None] - xyz[None], axis=-1)
    for q in range(0, 200, 17):
        r = np.linalg.norm(xyz - pts[q], axis=1)
        P = np.ones(len(els))
        for i in range(len(els)):
            for k in range(len(els)):
                if i != k:
                    mu = (r[i] - r[k]) / d[i, k]
                    for _ in range(3):
                        mu = 1.5 * mu - 0.5 * mu ** 3
                    P[i] *= 0.5 * (1 - mu)
        assert np.allclose(W[:, q], P / P.sum(), rtol=1e-12, atol=1e-15)


def test_product_grid_equals_oracle_grid():
    """grid.molecular_grid (numpy partition) against the oracle's mkgrid:
    Slater radii, int(nrad R) shells, tabulated Lebedev, becke0."""
    for nang in (302, 194):
        els, xyz = systems.water_cluster(2, seed=1)
        g = grid.molecular_grid(els, xyz, nrad=50, nang=nang, partition="numpy")
        go = oracle.molecular_grid(els, xyz, 50, nang, nthreads=4)
        assert g.shape == go.shape
        assert np.array_equal(g[:, :3], go[:, :3])
        assert np.abs(g[:, 3] - go[:, 3]).max() <= 1e-15 * go[:, 3].max()
    # shells per atom: int(NRAD R) before the 40-bohr cut
    assert grid.radial_shells(75, grid.RADIUS["O"]) == 85
    assert grid.radial_shells(75, grid.RADIUS["H"]) == 75
    assert abs(grid.RADIUS["H"] * systems.BOHR_A - 0.53) < 1e-15
