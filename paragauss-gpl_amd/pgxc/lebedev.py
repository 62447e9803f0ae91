"""Lebedev rules on the unit sphere, octahedral symmetry: 302 points (degree
29) and 194 points (degree 23), the two angular grids of the workloads
(ParaGauss selects them by NANG, modules/grid_module.f90:371-372,1698-1701,1753).

The orbit parameters below were obtained in this project by solving the
moment equations of each rule's octahedral orbit structure with Gauss-Newton
from random starts (solve()):
  302: a1 + a3 + 6 b + 2 c + 2 d   (6 + 8 + 144 + 48 + 96 points, 24 unknowns)
  194: a1 + a2 + a3 + 4 b + c + d  (6 + 12 + 8 + 96 + 24 + 48 points, 16 unknowns)
A rule of a given structure and degree is unique, so the values coincide with
the published Lebedev-Laikov rules; the tests check exactness on every
monomial up to each rule's degree directly.
Weights are normalised to 1 (the reference multiplies by 4 pi,
modules/grid_module.f90:426-433).

Orbit generators (all sign changes and coordinate permutations):
  a1 (1, 0, 0)      a2 (1, 1, 0)/sqrt(2)      a3 (1, 1, 1)/sqrt(3)
  b  (l, l, m), m = sqrt(1 - 2 l^2)
  c  (p, q, 0), q = sqrt(1 - p^2)
  d  (r, s, t), t = sqrt(1 - r^2 - s^2)
"""
import itertools

import numpy as np

# rule -> (a1 weight, a2 weight, a3 weight, [(l, w)], [(p, w)], [(r, s, w)])
RULES = {
    302: (0.0008545911723757858, 0.0, 0.0035991192850245053,
          [(0.3515640345569994, 0.003449788424309843),
           (0.7011766416089655, 0.003650045807678685),
           (0.4729054132581182, 0.003576729661742814),
           (0.6566329410219505, 0.0036048226014175656),
           (0.09618308522484725, 0.0023521014137145714),
           (0.22196452362904395, 0.003108953122426444)],
          [(0.26441528870662284, 0.00298234496317498),
           (0.5718955891877325, 0.0036008209322161617)],
          [(0.12335485325856449, 0.412772408316982, 0.003392312205003125),
           (0.5448677372582008, 0.25100347517710475, 0.003571540554271933)]),
    194: (0.0017823404472476709, 0.005716905949977161, 0.0055733831788491554,
          [(0.6712973442695228, 0.005608704082588363),
           (0.44469331787173644, 0.00551877146727368),
           (0.28924656275754085, 0.005158237711804964),
           (0.12993354476502236, 0.004106777028168322)],
          [(0.34577021976110467, 0.005051846064614635)],
          [(0.8360360154824651, 0.1590417105383454, 0.005530248916233241)]),
}
DEGREE = {302: 29, 194: 23}


def _images(v):
    out = set()
    for perm in itertools.permutations(v):
        for sg in itertools.product((1.0, -1.0), repeat=3):
            out.add(tuple(s * c if c != 0.0 else 0.0 for s, c in zip(sg, perm)))
    return sorted(out)


def _orbits(a1, a2, a3, bs, cs, ds):
    pts, w = [], []
    gens = [((1.0, 0.0, 0.0), a1), ((3 ** -0.5,) * 3, a3)]
    if a2:
        gens.insert(1, ((2 ** -0.5, 2 ** -0.5, 0.0), a2))
    gens += [((l, l, np.sqrt(1 - 2 * l * l)), wt) for l, wt in bs]
    gens += [((p, np.sqrt(1 - p * p), 0.0), wt) for p, wt in cs]
    gens += [((r, s, np.sqrt(max(0.0, 1 - r * r - s * s))), wt) for r, s, wt in ds]
    for gen, wt in gens:
        im = _images(gen)
        pts += im
        w += [wt] * len(im)
    return np.array(pts), np.array(w)


_CACHE = {}


def rule(n=302):
    """(points (n,3), weights (n,)) with sum(weights) = 1; n = 302 or 194."""
    if n not in RULES:
        raise ValueError("Lebedev rules provided: %s points" % sorted(RULES))
    if n not in _CACHE:
        _CACHE[n] = _orbits(*RULES[n])
    return _CACHE[n]


def sphere_moment(a, b, c):
    """Mean of x^a y^b z^c over the unit sphere (even a, b, c)."""
    from math import lgamma, exp, log, pi
    al, be, ga = (a + 1) / 2, (b + 1) / 2, (c + 1) / 2
    return exp(log(2) + lgamma(al) + lgamma(be) + lgamma(ga) - lgamma(al + be + ga)) / (4 * pi)


def solve(n=302, seed=0, trials=20000):
    """Re-derive the orbit parameters of rule n (minutes of CPU); returns
    them in the layout of RULES[n]."""
    from scipy.optimize import least_squares
    a1, a2, a3, B, Cs, Ds = RULES[n]
    nb, nc, nd, has_a2 = len(B), len(Cs), len(Ds), a2 != 0.0
    E = np.array([(a, b, d - a - b) for d in range(0, DEGREE[n] + 1, 2) for a in range(0, d + 1, 2)
                  for b in range(0, a + 1, 2) if 0 <= d - a - b <= b], dtype=float)
    target = np.array([sphere_moment(*e) for e in E])

    def unpack(x):
        o = 3
        bs = [(x[o + 2 * k], x[o + 1 + 2 * k]) for k in range(nb)]
        o += 2 * nb
        cs = [(x[o + 2 * k], x[o + 1 + 2 * k]) for k in range(nc)]
        o += 2 * nc
        ds = [(x[o + 3 * k], x[o + 1 + 3 * k], x[o + 2 + 3 * k]) for k in range(nd)]
        return x[0], x[1] if has_a2 else 0.0, x[2], bs, cs, ds

    def resid(x):
        P, w = _orbits(*unpack(x))
        M = np.prod(np.where(E[None] == 0, 1.0, np.abs(P[:, None, :]) ** E[None]), axis=2)
        return (w @ M - target) / target

    rng = np.random.default_rng(seed)
    lo = [0, 0, 0] + [1e-3, 0] * nb + [1e-3, 0] * nc + [1e-3, 1e-3, 0] * nd
    hi = [0.05, 0.05 if has_a2 else 1e-12, 0.05] + [0.7071, 0.05] * nb + [0.9999, 0.05] * nc + \
        [0.999, 0.7, 0.05] * nd
    for _ in range(trials):
        x0 = rng.uniform(np.array(lo), np.array(hi))
        try:
            sol = least_squares(resid, x0, bounds=(lo, hi), xtol=1e-15, ftol=1e-15, gtol=1e-15,
                                max_nfev=300)
        except ValueError:
            continue
        if np.max(np.abs(sol.fun)) < 1e-11:
            return unpack(sol.x)
    raise RuntimeError("no solution found")
