"""Pins the oracle's XC build (modules/xc_hamiltonian.f90:1280-1580) by
brute-force numpy evaluation of the same sums on tiny inputs, plus the
identities the reference's arithmetic must satisfy."""
import numpy as np
import pytest

import oracle
from pgxc import basis, grid, systems
from pgxc.xc import unpack


def tiny_system(n_points=700, seed=0):
    els, xyz = systems.h2o()
    b = basis.seeded_basis(els, xyz, seed=seed)
    g = grid.molecular_grid(els, xyz, nrad=20, nang=302)
    rng = np.random.default_rng(seed)
    g = g[np.sort(rng.choice(len(g), n_points, replace=False))]
    C, occ = systems.random_orbitals(b.nbas, 5, seed)
    return b, np.asfortranarray(g), C, occ


def brute_build(b, g, P, mask):
    phi, grad = oracle.basis_eval(b, g[:, :3])
    w = g[:, 3]
    X = P @ phi
    rho = np.sum(phi * X, axis=0)
    gr = 2 * np.einsum("fcp,fp->cp", grad, X)
    gam = np.sum(gr ** 2, axis=0)
    f, fr, fg = oracle.functional(mask, rho, gam)
    gga = bool(mask & ((1 << 10) | (1 << 18)))
    A = 0.5 * fr * w
    Wm = A * phi
    if gga:
        B = w * 2 * fg * gr
        Wm = Wm + np.einsum("cp,fcp->fp", B, grad)
    H = Wm @ phi.T
    V = H + H.T
    return V, np.sum(f * w), np.sum(rho * w)


@pytest.mark.parametrize("mask", [oracle.MASK_PBE, oracle.MASK_LDA])
def test_build_matches_brute_force(mask):
    b, g, C, occ = tiny_system()
    P = (C * occ) @ C.T
    ham, e, q = oracle.xc_build(b, g, mask, C=C, occ=occ, veclen=128)
    V, e_ref, q_ref = brute_build(b, g, P, mask)
    assert e == pytest.approx(e_ref, rel=1e-12)
    assert q == pytest.approx(q_ref, rel=1e-12)
    assert np.max(np.abs(unpack(ham, b.nbas) - V)) < 1e-12 * np.max(np.abs(V))


def test_density_forms_agree():
    b, g, C, occ = tiny_system(seed=3)
    occ = np.array([2.0, 2.0, 1.5, 1.0, 0.5])
    P = (C * occ) @ C.T
    h1, e1, q1 = oracle.xc_build(b, g, oracle.MASK_PBE, C=C, occ=occ)
    h2, e2, q2 = oracle.xc_build(b, g, oracle.MASK_PBE, P=P)
    assert e1 == pytest.approx(e2, rel=1e-13, abs=1e-14)
    assert q1 == pytest.approx(q2, rel=1e-13)
    assert np.max(np.abs(h1 - h2)) < 1e-12 * np.max(np.abs(h1))


def test_batch_length_and_threads_do_not_change_result():
    b, g, C, occ = tiny_system(seed=4)
    ref = oracle.xc_build(b, g, oracle.MASK_PBE, C=C, occ=occ, veclen=128)
    for veclen, nt in ((64, 1), (97, 3), (700, 2)):
        h, e, q = oracle.xc_build(b, g, oracle.MASK_PBE, C=C, occ=occ, veclen=veclen, nthreads=nt)
        assert e == pytest.approx(ref[1], rel=1e-13)
        assert np.max(np.abs(h - ref[0])) < 1e-13 * np.max(np.abs(ref[0]))


def test_batch_range_partition_sums_to_whole():
    b, g, C, occ = tiny_system(seed=6)
    whole = oracle.xc_build(b, g, oracle.MASK_PBE, C=C, occ=occ, veclen=100)
    a = oracle.xc_build(b, g, oracle.MASK_PBE, C=C, occ=occ, veclen=100, batch_range=(0, 3))
    c = oracle.xc_build(b, g, oracle.MASK_PBE, C=C, occ=occ, veclen=100, batch_range=(3, -1))
    assert a[1] + c[1] == pytest.approx(whole[1], rel=1e-13)
    assert np.allclose(a[0] + c[0], whole[0], rtol=1e-12, atol=1e-15)


def test_exc_is_variational_derivative():
    # dE/dP_jk = V_jk: the potential matrix is the derivative of E_xc
    b, g, C, occ = tiny_system(seed=8)
    P = (C * occ) @ C.T
    ham, e0, _ = oracle.xc_build(b, g, oracle.MASK_PBE, P=P)
    V = unpack(ham, b.nbas)
    rng = np.random.default_rng(1)
    D = rng.normal(size=P.shape)
    D = (D + D.T) * 1e-6
    ep = oracle.xc_build(b, g, oracle.MASK_PBE, P=P + D)[1]
    em = oracle.xc_build(b, g, oracle.MASK_PBE, P=P - D)[1]
    assert (ep - em) / 2 == pytest.approx(np.sum(V * D), rel=1e-6)


def test_bad_functional_mask_raises():
    b, g, C, occ = tiny_system()
    with pytest.raises(ValueError):
        oracle.xc_build(b, g, 1 << 2, C=C, occ=occ)
