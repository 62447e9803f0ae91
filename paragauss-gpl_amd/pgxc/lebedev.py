"""302-point Lebedev rule on the unit sphere (degree 29), octahedral symmetry.

The orbit parameters below were obtained in this project by solving the
moment equations of the octahedral orbit structure a1 + a3 + 6 b + 2 c + 2 d
(6 + 8 + 144 + 48 + 96 = 302 points, 24 unknowns = the 24 O_h-invariant
harmonics of even degree <= 28) with Gauss-Newton from random starts; see
solve().  A rule of this structure is unique, so the values coincide with the
published Lebedev-Laikov rule; tests check exactness on all monomials of
degree <= 29 directly.
Weights are normalised to 1 (the reference multiplies by 4 pi,
modules/grid_module.f90:426-433).

Orbit generators (all sign changes and coordinate permutations):
  a1 (1, 0, 0)      a3 (1, 1, 1)/sqrt(3)
  b  (l, l, m), m = sqrt(1 - 2 l^2)
  c  (p, q, 0), q = sqrt(1 - p^2)
  d  (r, s, t), t = sqrt(1 - r^2 - s^2)
"""
import itertools

import numpy as np

A1_W = 0.0008545911723757858
A3_W = 0.0035991192850245053
B_ORBITS = [  # (l, weight)
    (0.3515640345569994, 0.003449788424309843),
    (0.7011766416089655, 0.003650045807678685),
    (0.4729054132581182, 0.003576729661742814),
    (0.6566329410219505, 0.0036048226014175656),
    (0.09618308522484725, 0.0023521014137145714),
    (0.22196452362904395, 0.003108953122426444),
]
C_ORBITS = [  # (p, weight)
    (0.26441528870662284, 0.00298234496317498),
    (0.5718955891877325, 0.0036008209322161617),
]
D_ORBITS = [  # (r, s, weight)
    (0.12335485325856449, 0.412772408316982, 0.003392312205003125),
    (0.5448677372582008, 0.25100347517710475, 0.003571540554271933),
]


def _images(v):
    out = set()
    for perm in itertools.permutations(v):
        for sg in itertools.product((1.0, -1.0), repeat=3):
            out.add(tuple(s * c if c != 0.0 else 0.0 for s, c in zip(sg, perm)))
    return sorted(out)


def _orbits(a1, a3, bs, cs, ds):
    pts, w = [], []
    for gen, wt in [((1.0, 0.0, 0.0), a1), ((3 ** -0.5,) * 3, a3)] + \
            [((l, l, np.sqrt(1 - 2 * l * l)), wt) for l, wt in bs] + \
            [((p, np.sqrt(1 - p * p), 0.0), wt) for p, wt in cs] + \
            [((r, s, np.sqrt(1 - r * r - s * s)), wt) for r, s, wt in ds]:
        im = _images(gen)
        pts += im
        w += [wt] * len(im)
    return np.array(pts), np.array(w)


_CACHE = {}


def rule(n=302):
    """(points (n,3), weights (n,)) with sum(weights) = 1."""
    if n != 302:
        raise ValueError("only the 302-point rule is provided")
    if n not in _CACHE:
        _CACHE[n] = _orbits(A1_W, A3_W, B_ORBITS, C_ORBITS, D_ORBITS)
    return _CACHE[n]


def sphere_moment(a, b, c):
    """Mean of x^a y^b z^c over the unit sphere (even a, b, c)."""
    from math import lgamma, exp, log, pi
    al, be, ga = (a + 1) / 2, (b + 1) / 2, (c + 1) / 2
    return exp(log(2) + lgamma(al) + lgamma(be) + lgamma(ga) - lgamma(al + be + ga)) / (4 * pi)


def solve(seed=0, trials=20000):
    """Re-derive the orbit parameters (minutes of CPU); returns the 24 numbers
    in the order a1, a3, (l, w) x 6, (p, w) x 2, (r, s, w) x 2."""
    from scipy.optimize import least_squares
    E = np.array([(a, b, d - a - b) for d in range(0, 30, 2) for a in range(0, d + 1, 2)
                  for b in range(0, a + 1, 2) if 0 <= d - a - b <= b], dtype=float)
    target = np.array([sphere_moment(*e) for e in E])

    def resid(x):
        bs = [(x[2 + 2 * k], x[3 + 2 * k]) for k in range(6)]
        cs = [(x[14 + 2 * k], x[15 + 2 * k]) for k in range(2)]
        ds = [(x[18 + 3 * k], x[19 + 3 * k], x[20 + 3 * k]) for k in range(2)]
        P, w = _orbits(x[0], x[1], bs, cs, ds)
        M = np.prod(np.where(E[None] == 0, 1.0, np.abs(P[:, None, :]) ** E[None]), axis=2)
        return (w @ M - target) / target

    rng = np.random.default_rng(seed)
    lo = [0, 0] + [1e-3, 0] * 6 + [1e-3, 0] * 2 + [1e-3, 1e-3, 0] * 2
    hi = [0.05, 0.05] + [0.7071, 0.05] * 6 + [0.9999, 0.05] * 2 + [0.999, 0.7, 0.05] * 2
    for _ in range(trials):
        x0 = rng.uniform(np.array(lo) + 1e-3, np.array(hi) - 1e-3)
        try:
            sol = least_squares(resid, x0, bounds=(lo, hi), xtol=1e-15, ftol=1e-15, gtol=1e-15,
                                max_nfev=300)
        except ValueError:
            continue
        if np.max(np.abs(sol.fun)) < 1e-11:
            return sol.x
    raise RuntimeError("no solution found")
