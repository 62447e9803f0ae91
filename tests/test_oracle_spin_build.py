"""Pins the oracle's open-shell XC build (ispin = 2,
modules/xc_hamiltonian.f90:1428-1460 and :1555-1567): the unpolarised limit
against the closed-shell build, the potential of each spin as the derivative
of E_xc in that spin's density matrix, the two density forms against each
other, and the interleaved output against a numpy evaluation of the sums."""
import numpy as np
import pytest

import oracle
from pgxc.xc import unpack

from test_oracle_build import tiny_system


def spin_orbitals(b, seed):
    """Alpha and beta orbitals: the closed-shell set and a rotated copy with
    one orbital fewer occupied, so the two spins differ everywhere."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(b.nbas, 6)))
    Ca, Cb = Q[:, :5], Q[:, 1:6]
    occ = np.stack([np.array([1.0, 1.0, 1.0, 0.7, 0.4]), np.array([1.0, 1.0, 0.9, 0.5, 0.0])], axis=1)
    return np.stack([Ca, Cb], axis=2), occ


def split(ham, n):
    return unpack(ham[0::2], n), unpack(ham[1::2], n)


@pytest.mark.parametrize("mask", [oracle.MASK_PBE, oracle.MASK_LDA])
def test_unpolarised_limit_is_closed_shell(mask):
    b, g, C, occ = tiny_system(seed=2)
    P = (C * occ) @ C.T
    h1, e1, q1 = oracle.xc_build(b, g, mask, P=P)
    h2, e2, q2 = oracle.xc_build(b, g, mask, P=np.stack([P / 2, P / 2], axis=2), ispin=2)
    assert e2 == pytest.approx(e1, rel=1e-12)
    assert q2 == pytest.approx(q1, rel=1e-13)
    scale = np.max(np.abs(h1))
    assert np.max(np.abs(h2[0::2] - h1)) < 1e-11 * scale
    assert np.max(np.abs(h2[1::2] - h1)) < 1e-11 * scale


@pytest.mark.parametrize("mask", [oracle.MASK_PBE, oracle.MASK_LDA])
def test_spin_potentials_are_derivatives_of_exc(mask):
    # dE/dP^s_jk = V^s_jk for each spin separately: fixes the alpha / beta
    # positions of the output and the cross-spin gradient term
    b, g, _, _ = tiny_system(seed=8)
    C, occ = spin_orbitals(b, 8)
    P = np.stack([(C[:, :, s] * occ[:, s]) @ C[:, :, s].T for s in range(2)], axis=2)
    ham, e0, _ = oracle.xc_build(b, g, mask, P=P, ispin=2)
    V = split(ham, b.nbas)
    rng = np.random.default_rng(1)
    for s in range(2):
        D = rng.normal(size=P.shape[:2])
        D = (D + D.T) * 1e-6
        Pp, Pm = P.copy(), P.copy()
        Pp[:, :, s] += D
        Pm[:, :, s] -= D
        ep = oracle.xc_build(b, g, mask, P=Pp, ispin=2)[1]
        em = oracle.xc_build(b, g, mask, P=Pm, ispin=2)[1]
        assert (ep - em) / 2 == pytest.approx(np.sum(V[s] * D), rel=1e-6)
        assert abs((ep - em) / 2 - np.sum(V[1 - s] * D)) > 1e-3 * abs(np.sum(V[s] * D))


def test_spin_density_forms_agree():
    b, g, _, _ = tiny_system(seed=3)
    C, occ = spin_orbitals(b, 3)
    P = np.stack([(C[:, :, s] * occ[:, s]) @ C[:, :, s].T for s in range(2)], axis=2)
    h1, e1, q1 = oracle.xc_build(b, g, oracle.MASK_PBE, C=C, occ=occ, ispin=2)
    h2, e2, q2 = oracle.xc_build(b, g, oracle.MASK_PBE, P=P, ispin=2)
    assert e1 == pytest.approx(e2, rel=1e-13, abs=1e-14)
    assert q1 == pytest.approx(q2, rel=1e-13)
    assert np.max(np.abs(h1 - h2)) < 1e-12 * np.max(np.abs(h1))


def test_spin_build_matches_numpy_sums():
    b, g, _, _ = tiny_system(seed=5)
    C, occ = spin_orbitals(b, 5)
    P = np.stack([(C[:, :, s] * occ[:, s]) @ C[:, :, s].T for s in range(2)], axis=2)
    ham, e, q = oracle.xc_build(b, g, oracle.MASK_PBE, P=P, ispin=2, veclen=96, nthreads=3)
    phi, grad = oracle.basis_eval(b, g[:, :3])
    w = g[:, 3]
    X = [P[:, :, s] @ phi for s in range(2)]
    rho = [np.sum(phi * x, axis=0) for x in X]
    gr = [2 * np.einsum("fcp,fp->cp", grad, x) for x in X]
    gaa, gbb, gab = (np.sum(gr[0] * gr[0], 0), np.sum(gr[1] * gr[1], 0), np.sum(gr[0] * gr[1], 0))
    f, fr, fg = oracle.functional_spin(oracle.MASK_PBE, rho[0], rho[1], gaa, gbb, gab)
    assert e == pytest.approx(np.sum(f * w), rel=1e-12)
    assert q == pytest.approx(np.sum((rho[0] + rho[1]) * w), rel=1e-12)
    for s in range(2):
        B = w * (2 * fg[:, s] * gr[s] + fg[:, 2] * gr[1 - s])
        Wm = 0.5 * fr[:, s] * w * phi + np.einsum("cp,fcp->fp", B, grad)
        H = Wm @ phi.T
        Vs = H + H.T
        assert np.max(np.abs(split(ham, b.nbas)[s] - Vs)) < 1e-12 * np.max(np.abs(Vs))
