"""Nuclear gradient of E_xc on the GPU, basis-centre term F1
(modules/post_scf_module.f90:150-185; grad_xc with weight_grads off): the
grid and its weights are fixed, so F1 is exactly the derivative of E_xc with
respect to the basis centres.  Pinned by central differences of the FP64
oracle's E_xc with the atoms' basis functions moved and the grid held, for
LDA and PBE, closed and open shell, and a basis with f and g shells (the
Hessian recursion up to l = 4)."""
import dataclasses
import os

import numpy as np
import pytest

import oracle
from pgxc import XCBuilder, basis, grid, systems
from pgxc.xc import BF16X3, FP64, LDA, PBE

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)
H = 4e-5


def fd_gradient(b, g, mask, **dens):
    """Central differences at h and h/2, Richardson-extrapolated (error
    O(h^4)): the basis reaches exponents of 5000, whose functions vary on
    0.014 bohr, so a plain central difference at 1e-4 is off by 2e-4."""
    def central(A, d, h):
        e = []
        for s in (1, -1):
            xyz = np.array(b.atom_xyz, dtype=np.float64, copy=True)
            xyz[A, d] += s * h
            e.append(oracle.xc_build(dataclasses.replace(b, atom_xyz=xyz), g, mask,
                                     nthreads=NT, **dens)[1])
        return (e[0] - e[1]) / (2 * h)
    out = np.zeros((b.n_atoms, 3))
    for A in range(b.n_atoms):
        for d in range(3):
            out[A, d] = (4 * central(A, d, H / 2) - central(A, d, H)) / 3
    return out


def check(gpu, fd):
    err = np.max(np.abs(gpu - fd))
    print("max|g|=%.3e max|g - fd|=%.3e" % (np.max(np.abs(fd)), err), flush=True)
    assert err <= 2e-6 * np.max(np.abs(fd)) + 1e-7


@pytest.fixture(scope="module")
def water():
    els, xyz = systems.h2o()
    b = basis.seeded_basis(els, xyz, seed=1, shell_set="small")
    g = grid.molecular_grid(els, xyz, nrad=30, nang=302)
    C, occ = systems.random_orbitals(b.nbas, 5, seed=2)
    return b, g, C, occ


@pytest.mark.parametrize("functional,mask", [(PBE, oracle.MASK_PBE), (LDA, oracle.MASK_LDA)],
                         ids=["PBE", "LDA"])
def test_gradient_matches_finite_differences(water, functional, mask):
    b, g, C, occ = water
    with XCBuilder(b, g) as xc:
        xc.set_orbitals(C, occ)
        xc.build(functional, FP64)
        gr = xc.gradient()
    check(gr, fd_gradient(b, g, mask, C=C, occ=occ))


def test_gradient_open_shell(water):
    b, g, _, _ = water
    rng = np.random.default_rng(4)
    Q, _ = np.linalg.qr(rng.normal(size=(b.nbas, 6)))
    C = np.stack([Q[:, :5], Q[:, 1:6]], axis=2)
    occ = np.stack([np.ones(5), np.array([1.0, 1.0, 0.8, 0.5, 0.0])], axis=1)
    with XCBuilder(b, g, chunk_points=8192) as xc:
        xc.set_orbitals(C, occ)
        xc.build(PBE, FP64)
        gr = xc.gradient()
    check(gr, fd_gradient(b, g, oracle.MASK_PBE, C=C, occ=occ, ispin=2))


def test_gradient_f_and_g_shells():
    els, xyz = systems.h2o()
    b = basis.seeded_basis(els, xyz, seed=7, shell_set="spdfg", exp_range=(0.1, 500.0))
    g = grid.molecular_grid(els, xyz, nrad=24, nang=302)
    C, occ = systems.random_orbitals(b.nbas, 5, seed=8)
    with XCBuilder(b, g) as xc:
        xc.set_orbitals(C, occ)
        xc.build(PBE, FP64)
        gr = xc.gradient()
        # after a split-BF16 full build, the gradient from its coefficients
        # carries that build's error: the mixed full-build bar of V, 10 x 5e-5
        # relative
        xc.build(PBE, BF16X3)
        gm = xc.gradient()
    fd = fd_gradient(b, g, oracle.MASK_PBE, C=C, occ=occ)
    check(gr, fd)
    assert np.max(np.abs(gm - gr)) <= 5e-4 * np.max(np.abs(gr))


def test_gradient_and_energy_need_a_build(water):
    """The gradient reads the last build's coefficients and the energy the
    current density: before them both fail with the library's state error."""
    from pgxc import _lib
    b, g, C, occ = water
    with XCBuilder(b, g) as xc:
        with pytest.raises(_lib.PgxcError):
            xc.energy(PBE, FP64)
        xc.set_orbitals(C, occ)
        with pytest.raises(_lib.PgxcError):
            xc.gradient()
        terms, q = xc.energy(PBE, FP64)  # the density suffices for energies
        assert q > 0
        xc.build(PBE, FP64)
        bad = np.zeros(g.shape[0], dtype=np.int32)
        bad[7] = b.n_atoms  # an atom index out of range
        with pytest.raises(_lib.PgxcError):
            xc.gradient_grid(bad)
        with pytest.raises(ValueError):
            xc.gradient_grid(bad[:-1])


def water_grid(xyz, nrad=30, nang=302):
    els, _ = systems.h2o()
    return grid.molecular_grid(els, xyz, nrad=nrad, nang=nang, partition="numpy")


@pytest.mark.parametrize("functional,mask,ispin", [(PBE, oracle.MASK_PBE, 1),
                                                   (LDA, oracle.MASK_LDA, 1),
                                                   (PBE, oracle.MASK_PBE, 2)],
                         ids=["PBE", "LDA", "PBE-open-shell"])
def test_gradient_with_moving_grid(functional, mask, ispin):
    """weight_grads on: every atom's grid moves with it and the Becke weights
    follow.  Against Richardson-extrapolated central differences of the
    oracle's E_xc on grids regenerated at the displaced geometry; and the sum
    over atoms vanishes (a rigid translation leaves E_xc unchanged)."""
    els, xyz = systems.h2o()
    b = basis.seeded_basis(els, xyz, seed=1, shell_set="small")
    g = water_grid(xyz)
    # atoms' grids in order, each truncated where the radial grid ends
    sizes = [len(grid.atomic_grid(c, 30, 302, grid.RADIUS[e])[0]) for e, c in zip(els, xyz)]
    owner = np.repeat(np.arange(len(els)), sizes)
    assert owner.shape[0] == g.shape[0]
    C, occ = systems.random_orbitals(b.nbas, 5, seed=2)
    if ispin == 2:
        rng = np.random.default_rng(4)
        Q, _ = np.linalg.qr(rng.normal(size=(b.nbas, 6)))
        C = np.stack([Q[:, :5], Q[:, 1:6]], axis=2)
        occ = np.stack([np.ones(5), np.array([1.0, 1.0, 0.8, 0.5, 0.0])], axis=1)
    dens = dict(C=C, occ=occ) if ispin == 1 else dict(C=C, occ=occ, ispin=2)
    with XCBuilder(b, g, chunk_points=8192) as xc:
        xc.set_orbitals(C, occ)
        xc.build(functional, FP64)
        gr = xc.gradient_grid(owner)

    def central(A, d, h):
        e = []
        for s in (1, -1):
            x2 = np.array(xyz, dtype=np.float64, copy=True)
            x2[A, d] += s * h
            e.append(oracle.xc_build(dataclasses.replace(b, atom_xyz=x2), water_grid(x2), mask,
                                     nthreads=NT, **dens)[1])
        return (e[0] - e[1]) / (2 * h)
    fd = np.zeros_like(gr)
    for A in range(len(els)):
        for d in range(3):
            fd[A, d] = (4 * central(A, d, H / 2) - central(A, d, H)) / 3
    print("sum over atoms", gr.sum(0), flush=True)
    check(gr, fd)
    assert np.max(np.abs(gr.sum(0))) <= 1e-7 * np.max(np.abs(gr)) + 1e-9
