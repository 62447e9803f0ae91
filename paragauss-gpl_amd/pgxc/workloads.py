"""The XC-build workloads, from seeds: what the measurements time and the
size tests check.  Each returns (elements, basis, grid, C, occ) with the grid
Becke-partitioned on the GPU where it is large.

  h2o            H2O, small basis (24 functions), 75 x 302 grid, 5 occupied
  h2o32          (H2O)32 on a jittered lattice, medium basis (1056 functions),
                 75 x 302 grid (~1.96 M points), 160 occupied
  batch64        64 seeded C/H/N/O molecules of 3-60 atoms, small basis,
                 50-75 x 194/302 grids, nbas/5 occupied (a list of workloads)
  water_solute   (H2O)128 + a 40-atom C/H/N/O solute, medium basis (4982
                 functions), 72 x 302 grid (~8 M points), 640 occupied
"""
import numpy as np

from . import basis, grid, systems


def h2o(partition="auto"):
    els, xyz = systems.h2o()
    b = basis.seeded_basis(els, xyz, seed=1, shell_set="small")
    g = grid.molecular_grid(els, xyz, nrad=75, nang=302, partition=partition)
    C, occ = systems.random_orbitals(b.nbas, 5, seed=13)
    return els, b, g, C, occ


def h2o32(partition="gpu"):
    els, xyz = systems.water_cluster(32, seed=11)
    b = basis.seeded_basis(els, xyz, seed=12, shell_set="medium")
    g = grid.molecular_grid(els, xyz, nrad=75, nang=302, partition=partition)
    C, occ = systems.random_orbitals(b.nbas, 160, seed=13)
    return els, b, g, C, occ


def batch64_specs():
    """(atoms, nrad, nang) of the 64 molecules, in order."""
    rng = np.random.default_rng(2024)
    out = []
    for _ in range(64):
        na = int(rng.integers(3, 61))
        nrad = int(rng.integers(50, 76))
        nang = int(rng.choice([194, 302]))
        out.append((na, nrad, nang))
    return out


def batch64_molecule(i, partition="gpu"):
    na, nrad, nang = batch64_specs()[i]
    els, xyz = systems.random_molecule(na, seed=1000 + i)
    b = basis.seeded_basis(els, xyz, seed=2000 + i, shell_set="small")
    g = grid.molecular_grid(els, xyz, nrad=nrad, nang=nang, partition=partition)
    C, occ = systems.random_orbitals(b.nbas, max(1, b.nbas // 5), seed=3000 + i)
    return els, b, g, C, occ


def batch64(partition="gpu"):
    return [batch64_molecule(i, partition) for i in range(64)]


def water_solute(partition="gpu"):
    els, xyz = systems.water_solute(128, 40, seed=41)
    b = basis.seeded_basis(els, xyz, seed=42, shell_set="medium")
    g = grid.molecular_grid(els, xyz, nrad=72, nang=302, partition=partition)
    C, occ = systems.random_orbitals(b.nbas, 640, seed=43)
    return els, b, g, C, occ
