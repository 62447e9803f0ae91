"""Seeded molecular systems and densities for the XC-build workloads.

Lengths are converted with ParaGauss's own constant, 1 bohr = 0.52917706 A
(modules/constants.f90:107-108).
"""
import numpy as np

BOHR_A = 0.52917706
ANG = 1.0 / BOHR_A  # bohr per angstrom


def h2o(center=(0.0, 0.0, 0.0), rot=None):
    """O at the origin, H at 0.9572 A, HOH angle 104.52 deg (bohr)."""
    r = 0.9572 * ANG
    th = np.deg2rad(104.52) / 2
    xyz = np.array([[0.0, 0.0, 0.0],
                    [r * np.sin(th), 0.0, r * np.cos(th)],
                    [-r * np.sin(th), 0.0, r * np.cos(th)]])
    if rot is not None:
        xyz = xyz @ rot.T
    return ["O", "H", "H"], xyz + np.asarray(center)


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def water_cluster(n, seed, spacing_a=3.1, jitter_a=0.25, min_oo_a=2.7):
    """(H2O)_n on a jittered cubic lattice, O-O >= min_oo_a, random orientations."""
    rng = np.random.default_rng(seed)
    side = int(np.ceil(n ** (1.0 / 3.0)))
    sites = np.array([(i, j, k) for i in range(side) for j in range(side) for k in range(side)],
                     dtype=float)[:n] * spacing_a
    while True:
        o = sites + rng.uniform(-jitter_a, jitter_a, size=sites.shape)
        d = np.linalg.norm(o[:, None] - o[None], axis=-1) + np.eye(n) * 1e9
        if d.min() >= min_oo_a:
            break
    els, xyz = [], []
    for i in range(n):
        e, x = h2o(o[i] * ANG, random_rotation(rng))
        els += e
        xyz.append(x)
    xyz = np.vstack(xyz)
    return els, xyz - xyz.mean(axis=0)


def random_molecule(natoms, seed, bond=(1.0, 1.5), min_nonbonded=2.0):
    """Tree-grown C/H/N/O molecule: bonds in [1.0, 1.5] A, every pair that is
    not bonded at least min_nonbonded A apart."""
    rng = np.random.default_rng(seed)
    els = [rng.choice(["C", "N", "O"])]
    xyz = [np.zeros(3)]
    bonds = set()
    while len(xyz) < natoms:
        for _ in range(2000):
            parent = int(rng.integers(len(xyz)))
            v = rng.normal(size=3)
            v /= np.linalg.norm(v)
            cand = xyz[parent] + v * rng.uniform(*bond)
            ok = True
            for j, x in enumerate(xyz):
                if j != parent and np.linalg.norm(cand - x) < min_nonbonded:
                    ok = False
                    break
            if ok:
                bonds.add((parent, len(xyz)))
                xyz.append(cand)
                els.append(rng.choice(["C", "H", "N", "O"], p=[0.35, 0.35, 0.15, 0.15]))
                break
        else:
            raise RuntimeError("random_molecule: could not place atom %d" % len(xyz))
    return els, np.array(xyz) * ANG


def water_solute(n_water, n_solute, seed, gap_a=3.0, min_dist_a=2.0):
    """(H2O)_n_water cluster with a seeded n_solute-atom C/H/N/O molecule placed
    beside it along +x: the solute's nearest atom gap_a beyond the cluster's
    extent, every solute-water distance at least min_dist_a (bohr)."""
    ew, xw = water_cluster(n_water, seed=seed)
    es, xs = random_molecule(n_solute, seed=seed + 1)
    xs = xs - xs.mean(axis=0)
    shift = xw[:, 0].max() - xs[:, 0].min() + gap_a * ANG
    xs = xs + np.array([shift, 0.0, 0.0])
    d = np.linalg.norm(xs[:, None] - xw[None], axis=-1)
    assert d.min() >= min_dist_a * ANG
    return list(ew) + list(es), np.vstack([xw, xs])


def random_orbitals(nbas, nocc, seed):
    """QR-orthonormalised N(0,1) coefficients (nbas x nocc) and occ = 1."""
    rng = np.random.default_rng(seed)
    C = np.linalg.qr(rng.normal(size=(nbas, nocc)))[0]
    return np.asfortranarray(C), np.ones(nocc)


def density_from_orbitals(C, occ):
    return (C * occ) @ C.T


def perturbed_orbitals(C, norm_f, seed):
    """Orthonormal C' = qr(C + e X) with e chosen so that ||C'C'^T - C C^T||_F
    equals norm_f (to 1 %): a positive semidefinite density close to C C^T."""
    rng = np.random.default_rng(seed)
    X = rng.normal(size=C.shape)
    P0 = C @ C.T

    def dist(e):
        Q = np.linalg.qr(C + e * X)[0]
        Q = Q * np.sign(np.sum(Q * C, axis=0))
        return Q, np.linalg.norm(Q @ Q.T - P0)

    e = norm_f / max(dist(1e-3)[1] / 1e-3, 1e-300)
    for _ in range(6):
        Q, d = dist(e)
        if abs(d - norm_f) <= 0.01 * norm_f:
            break
        e *= norm_f / d
    return np.asfortranarray(Q)
