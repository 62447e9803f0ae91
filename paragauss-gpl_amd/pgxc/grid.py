"""Becke-partitioned atomic integration grids (input of the XC build).

Follows ParaGauss's grid construction (grid_main_calc,
modules/grid_module.f90:357-504, with the SCF defaults: Slater radii,
no pruning, no rotation, no cell-size adjustment since weight_grads = T):
  radius   Slater radius of the element (rslater_f, :1029-1079; angstrom
           times the reference's angstrom = 1/bohr, modules/constants.f90:108)
  shells   NRADPTS = int(NRAD * radius) (SET_GRID_PAR, :2819)
  radial   Pople/Gill grid r_i = R (i / (n+1-i))^2, w_i = 2 R^3 (n+1) i^5 / (n+1-i)^7,
           n = NRADPTS, shells with r < 40 bohr (RADIALGRID, :1246-1301)
  angular  Lebedev rule (302 points, degree 29, or 194, degree 23) from the
           reference's tabulated parameters (pgxc/lebedev.py)
  product  w = w_ang * w_rad * 4 pi  (:426-433), shell by shell
  Becke    partition of every atomic grid (becke0, :2442-2497; mkgrid,
           :3045-3164), on the GPU through pgxc_becke_weights, or in numpy
           for small systems
  layout   atoms concatenated in order, the linear grid grdpts(npts, 4)
           (:3045-3164), in which batches are contiguous slices.
"""
import numpy as np

from . import lebedev
from .systems import ANG

RAD_LIMIT = 40.0
# Slater radii in angstrom of the elements the workloads use
# (modules/grid_module.f90:1051-1062); RADIUS in bohr.
SLATER_A = {"H": 0.53, "C": 0.70, "N": 0.65, "O": 0.60}
RADIUS = {e: r * ANG for e, r in SLATER_A.items()}


def radial_shells(nrad, R):
    """NRADPTS of an atom: int(NRAD * radius)."""
    return int(nrad * R)


def radial_grid(n, R):
    """(r, w) of the Pople radial grid of n points, the shells inside the
    40-bohr limit, innermost first."""
    n1 = float(n + 1)
    r, w = [], []
    for i in range(1, n + 1):
        h = n1 - i
        x = i / h
        ri = R * (x * x)
        if ri >= RAD_LIMIT:
            break
        x2 = x * x
        r.append(ri)
        w.append(2 * (R * R * R) * n1 * (x2 * x2 * x) / (h * h))
    return np.array(r), np.array(w)


def atomic_grid(center, nrad, nang, R):
    """One-centre grid (points, weights) ordered radial shell by shell:
    NRAD = nrad, radius R (bohr), so int(nrad R) Pople shells."""
    ang, wang = lebedev.rule(nang)
    r, wr = radial_grid(radial_shells(nrad, R), R)
    pts = (r[:, None, None] * ang[None, :, :]).reshape(-1, 3) + np.asarray(center)
    w = (wr[:, None] * wang[None, :] * (16.0 * np.arctan(1.0))).reshape(-1)
    return pts, w


def becke_numpy(atom_xyz, pts, owner, w_atomic):
    """Becke partition in numpy (small systems; same formula as the GPU kernel)."""
    A = np.asarray(atom_xyz)
    na = A.shape[0]
    r = np.linalg.norm(pts[:, None, :] - A[None, :, :], axis=-1)  # (npts, na)
    P = np.ones((pts.shape[0], na))
    for i in range(na):
        for k in range(na):
            if i == k:
                continue
            mu = (r[:, i] - r[:, k]) / np.linalg.norm(A[i] - A[k])
            for _ in range(3):
                mu = mu * (3.0 - mu * mu) / 2.0
            P[:, i] *= (1.0 - mu) / 2.0
    return w_atomic * P[np.arange(len(owner)), owner] / P.sum(axis=1)


def owners(elements, nrad=75, nang=302):
    """Atom of every point of molecular_grid's grid (the atomic grid holding it)."""
    sizes = [radial_grid(radial_shells(nrad, RADIUS[e]), RADIUS[e])[0].size * nang
             for e in elements]
    return np.repeat(np.arange(len(elements), dtype=np.int32), sizes)


def molecular_grid(elements, atom_xyz, nrad=75, nang=302, partition="auto",
                   drop_zero=False):
    """Linear grid (npts, 4) = x, y, z, w with atoms concatenated in order."""
    pts, wat, own = [], [], []
    for a, (el, c) in enumerate(zip(elements, atom_xyz)):
        p, w = atomic_grid(c, nrad, nang, RADIUS[el])
        pts.append(p)
        wat.append(w)
        own.append(np.full(len(w), a, dtype=np.int32))
    pts = np.vstack(pts)
    wat = np.concatenate(wat)
    own = np.concatenate(own)
    if partition == "auto":
        partition = "numpy" if pts.shape[0] * len(elements) ** 2 < 5e7 else "gpu"
    if len(elements) == 1 or partition == "none":
        w = wat
    elif partition == "numpy":
        w = becke_numpy(atom_xyz, pts, own, wat)
    else:
        from .xc import becke_weights
        w = becke_weights(atom_xyz, pts, own, wat)
    grid = np.column_stack([pts, w])
    if drop_zero:
        grid = grid[w != 0.0]
    return np.asfortranarray(grid)
