"""Lebedev rules on the unit sphere, octahedral symmetry: 302 points (degree
29) and 194 points (degree 23), the two angular grids of the workloads
(ParaGauss selects them by NANG = 291 and 231, modules/grid_module.f90:
371-372,1698-1701,1753).

rule(n) is the grid the reference integrates on: the orbit parameters and
weights it tabulates to 12 digits (INIGRD, modules/grid_module.f90:1700-1814),
expanded into points orbit by orbit in its order (a1, a2, a3, b_k, c_k, d_k;
GETA1/GETA2/GETA3/GETBK/GETC1/GETDK, :1826-2098).  The tabulated digits are
part of the integration grid, so E_xc on this grid equals the reference's
only with the same digits: they differ from the exact rule by up to 6e-12
relative, which moves E_xc of (H2O)32 by ~5e-9 Eh.

rule(n, "solved") is the exact rule of the same orbit structure, solved in
this project from the moment equations with Gauss-Newton from random starts
(solve()); a rule of a given structure and degree is unique, so it is the
published Lebedev-Laikov rule, and it pins the tabulated one (the tests
check both):
  302: a1 + a3 + 6 b + 2 c + 2 d   (6 + 8 + 144 + 48 + 96 points, 24 unknowns)
  194: a1 + a2 + a3 + 4 b + c + d  (6 + 12 + 8 + 96 + 24 + 48 points, 16 unknowns)
Weights are normalised to 1 (the reference multiplies by 4 pi,
modules/grid_module.f90:426-433).

Orbit generators (all sign changes and coordinate permutations):
  a1 (1, 0, 0)      a2 (1, 1, 0)/sqrt(2)      a3 (1, 1, 1)/sqrt(3)
  b  (l, l, m), l = sqrt((1 - m^2)/2)
  c  (p, q, 0), p, q = sqrt((1 +- sqrt(1 - 4 v))/2), v = p^2 q^2
  d  (u, w, r), r = sqrt(1 - u^2 - w^2)
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


# The reference's tabulated rules: weights (a1, a2, a3), [(m, w_b)], [(v, w_c)],
# [(u, w, w_d)] (INIGRD, modules/grid_module.f90:1711-1729 and 1760-1785).
TABLES = {
    194: (1.78234044724E-03, 5.71690594998E-03, 5.57338317884E-03,
          [(0.777493219315, 5.51877146727E-03), (0.912509096867, 5.15823771181E-03),
           (0.314196994183, 5.60870408259E-03), (0.982972302707, 4.10677702817E-03)],
          [(2.00 / 19.0, 5.05184606462E-03)],
          [(0.159041710538, 0.525118572443, 5.53024891623E-03)]),
    302: (8.54591172878E-04, 0.0, 3.59911928502E-03,
          [(0.129238672710, 3.65004580768E-03), (0.371034178385, 3.60482260142E-03),
           (0.743452042987, 3.57672966173E-03), (0.867643624544, 3.44978842429E-03),
           (0.949454317226, 3.10895312238E-03), (0.990705621379, 2.35210141366E-03)],
          [(0.220093335298, 3.60082093222E-03), (0.650272754659E-01, 2.98234496317E-03)],
          [(0.800072749407, 0.544867737258, 3.57154055427E-03),
           (0.902442529533, 0.412772408317, 3.39231220501E-03)]),
}

_SIGNS = [(i, j, k) for i in (-1.0, 1.0) for j in (-1.0, 1.0) for k in (-1.0, 1.0)]


def _tab_a1():
    return [(0.0, 0.0, 1.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), (0.0, -1.0, 0.0),
            (1.0, 0.0, 0.0), (-1.0, 0.0, 0.0)]


def _tab_a2():
    s = 1.0 / np.sqrt(2.0)
    out = []
    for a, b in ((0, 1), (0, 2), (1, 2)):  # the (x, y), (x, z), (y, z) planes
        for sa, sb in ((s, s), (s, -s), (-s, s), (-s, -s)):
            v = [0.0, 0.0, 0.0]
            v[a], v[b] = sa, sb
            out.append(tuple(v))
    return out


def _tab_a3():
    t = 1.0 / np.sqrt(3.0)
    return [(i * t, j * t, k * t) for i, j, k in _SIGNS]


def _tab_b(m):
    l = np.sqrt((1.0 - m * m) / 2.0)
    return ([(i * l, j * l, k * m) for i, j, k in _SIGNS] +
            [(i * l, j * m, k * l) for i, j, k in _SIGNS] +
            [(i * m, j * l, k * l) for i, j, k in _SIGNS])


def _tab_c(v):
    d = np.sqrt(1.0 - 4.0 * v)
    p, q = np.sqrt((1.0 + d) / 2.0), np.sqrt((1.0 - d) / 2.0)
    out = []
    for a, b in ((p, q), (q, p)):  # (p, q) first, then the swapped pair
        for i, k in ((0, 1), (0, 2), (1, 2)):
            for sa, sb in ((1.0, 1.0), (1.0, -1.0), (-1.0, 1.0), (-1.0, -1.0)):
                x = [0.0, 0.0, 0.0]
                x[i], x[k] = sa * a, sb * b
                out.append(tuple(x))
    return out


def _tab_d(u, w):
    r = np.sqrt(1.0 - u * u - w * w)
    first = ([(i * u, j * w, k * r) for i, j, k in _SIGNS] +
             [(i * w, j * r, k * u) for i, j, k in _SIGNS] +
             [(i * r, j * u, k * w) for i, j, k in _SIGNS])
    return first + [(x, z, y) for x, y, z in first]


def _table_rule(n):
    a1, a2, a3, bs, cs, ds = TABLES[n]
    orbits = [(_tab_a1(), a1)]
    if a2:
        orbits.append((_tab_a2(), a2))
    orbits.append((_tab_a3(), a3))
    orbits += [(_tab_b(m), wt) for m, wt in bs]
    orbits += [(_tab_c(v), wt) for v, wt in cs]
    orbits += [(_tab_d(u, w), wt) for u, w, wt in ds]
    pts = [p for o, _ in orbits for p in o]
    wts = [wt for o, wt in orbits for _ in o]
    return np.array(pts), np.array(wts)


_CACHE = {}


def rule(n=302, source="table"):
    """(points (n,3), weights (n,)) with sum(weights) = 1 (to the table's
    digits); n = 302 or 194.  source: "table" (the reference's grid) or
    "solved" (the exact rule)."""
    if n not in RULES:
        raise ValueError("Lebedev rules provided: %s points" % sorted(RULES))
    if source not in ("table", "solved"):
        raise ValueError("source must be 'table' or 'solved'")
    key = (n, source)
    if key not in _CACHE:
        _CACHE[key] = _table_rule(n) if source == "table" else _orbits(*RULES[n])
    return _CACHE[key]


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
