"""Becke-partitioned atomic integration grids (input of the XC build).

Follows ParaGauss's grid construction (modules/grid_module.f90:357-504):
  radial   Pople/Gill grid r_i = R (i / (n+1-i))^2, w_i = 2 R^3 (n+1) i^5 / (n+1-i)^7,
           truncated at 40 bohr (RADIALGRID, :1246-1301)
  angular  Lebedev rule (302 points, degree 29, or 194, degree 23; own
           solutions of the octahedral moment equations, pgxc/lebedev.py)
  product  w = w_ang * w_rad * 4 pi  (:426-433)
  Becke    partition of every atomic grid (becke0, :2442-2497), on the GPU
           through pgxc_becke_weights, or in numpy for small systems
  layout   atoms concatenated in order, the linear grid grdpts(npts, 4)
           (:3045-3164), in which batches are contiguous slices.
"""
import numpy as np

from . import lebedev

RAD_LIMIT = 40.0
# Pople-grid radius parameter per element (bohr); these set where the
# radial points cluster, any positive value gives a valid grid.
RADIUS = {"H": 0.661, "C": 1.323, "N": 1.228, "O": 1.134}


def radial_grid(nrad, R):
    """(r, w) of the truncated Pople radial grid, innermost first."""
    n1 = nrad + 1
    r, w = [], []
    for i in range(1, nrad + 1):
        ri = R * (i / (n1 - i)) ** 2
        if ri >= RAD_LIMIT:
            break
        r.append(ri)
        w.append(2 * R ** 3 * n1 * (i / (n1 - i)) ** 5 / (n1 - i) ** 2)
    return np.array(r), np.array(w)


def atomic_grid(center, nrad, nang, R):
    """One-centre grid (points, weights) ordered radial shell by shell."""
    ang, wang = lebedev.rule(nang)
    r, wr = radial_grid(nrad, R)
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
