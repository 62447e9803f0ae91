"""Pins the oracle's basis evaluation: solid harmonics (reference build +
golden vectors + closed forms + identities), primitive norms (unit
self-overlap), and the assembled basis values (brute force)."""
import os

import numpy as np
import pytest

import oracle
from pgxc import basis, systems

GOLD = os.path.join(os.path.dirname(__file__), "golden", "harmonics.npz")


def test_harmonics_golden_vectors():
    d = np.load(GOLD)
    C, G = oracle.solid_harmonics(4, d["pts"] - d["center"])
    assert np.max(np.abs(C - d["C"])) < 1e-12 * np.max(np.abs(d["C"]))
    assert np.max(np.abs(G - d["G"])) < 1e-12 * np.max(np.abs(d["G"]))


@pytest.mark.skipif(oracle.ref_lib() is None, reason="oracle/_ref not built (no reference tree)")
def test_harmonics_match_reference_build():
    rng = np.random.default_rng(11)
    pts = rng.normal(size=(40, 3)) * 2
    ctr = rng.normal(size=3)
    C, G = oracle.solid_harmonics(4, pts - ctr)
    Cr, Gr = oracle.ref_solid_harmonics(4, pts, ctr)
    assert np.max(np.abs(C - Cr)) <= 1e-13 * np.max(np.abs(Cr))
    assert np.max(np.abs(G - Gr)) <= 1e-13 * np.max(np.abs(Gr))


def test_harmonics_closed_forms_l2():
    rng = np.random.default_rng(1)
    x, y, z = rng.normal(size=(3, 30))
    C, _ = oracle.solid_harmonics(2, np.column_stack([x, y, z]))
    r2 = x * x + y * y + z * z
    s3 = np.sqrt(3.0)
    want = [np.ones_like(x), z, x, y, (3 * z * z - r2) / 2, s3 * z * x, s3 * z * y,
            s3 * (x * x - y * y) / 2, s3 * x * y]
    for i, w in enumerate(want):
        assert np.allclose(C[i], w, rtol=1e-14, atol=1e-14)


def test_harmonics_racah_norm_euler_and_fd():
    rng = np.random.default_rng(2)
    xyz = rng.normal(size=(25, 3))
    L = 4
    C, G = oracle.solid_harmonics(L, xyz)
    r2 = np.sum(xyz ** 2, axis=1)
    for l in range(L + 1):
        blk = C[l * l:(l + 1) ** 2]
        assert np.allclose(np.sum(blk ** 2, axis=0), r2 ** l, rtol=1e-13)  # Racah
        euler = np.einsum("mcp,pc->mp", G[l * l:(l + 1) ** 2], xyz)
        assert np.allclose(euler, l * blk, rtol=1e-12, atol=1e-12)  # r.grad C = l C
    h = 1e-6
    for c in range(3):
        d = np.zeros(3)
        d[c] = h
        Cp, _ = oracle.solid_harmonics(L, xyz + d)
        Cm, _ = oracle.solid_harmonics(L, xyz - d)
        assert np.allclose((Cp - Cm) / (2 * h), G[:, c, :], rtol=1e-7, atol=1e-7)


@pytest.mark.parametrize("L", [0, 1, 2, 3])
def test_primitive_unit_self_overlap(L):
    # int (N C_lm e^{-a r^2})^2 d^3r = N^2 4pi/(2l+1) int r^(2l+2) e^{-2 a r^2} dr = 1
    from scipy.integrate import quad
    for a in (0.1, 1.3, 47.0, 5000.0):
        N = oracle.prim_norm(L, a)
        assert N == pytest.approx(basis.legacy_norm(L, a), rel=1e-15)
        rad = quad(lambda r: r ** (2 * L + 2) * np.exp(-2 * a * r * r), 0, np.inf,
                   epsabs=0, epsrel=1e-13)[0]
        assert N * N * 4 * np.pi / (2 * L + 1) * rad == pytest.approx(1.0, rel=1e-10)


def brute_basis(b, pts):
    """Independent evaluation of every function from its definition."""
    out, grads = [], []
    for s in b.shells:
        R = b.atom_xyz[s.atom]
        rel = pts - R
        r2 = np.sum(rel ** 2, axis=1)
        C, G = oracle.solid_harmonics(s.l, rel)
        prim = np.where(np.outer(r2, s.exps) < 50.0,
                        s.norms[None] * np.exp(-np.outer(r2, s.exps)), 0.0)   # (p, e)
        dprim = prim * (-2 * s.exps[None])
        rad = np.concatenate([prim[:, :s.n_unc], prim @ s.contr], axis=1)
        drad = np.concatenate([dprim[:, :s.n_unc], dprim @ s.contr], axis=1)
        for m in range(2 * s.l + 1):
            cm = C[s.l * s.l + m]
            gm = G[s.l * s.l + m]
            for k in range(rad.shape[1]):
                out.append(cm * rad[:, k])
                grads.append(rel.T * (cm * drad[:, k]) + gm * rad[:, k])
    return np.array(out), np.array(grads)


def test_basis_eval_brute_force():
    els, xyz = systems.h2o()
    b = basis.seeded_basis(els, xyz, seed=5)
    rng = np.random.default_rng(4)
    pts = rng.normal(size=(200, 3)) * 2.0
    pts[:3] = xyz  # on the nuclei
    pts[3] = [40.0, 0, 0]  # everything cut off
    phi, grad = oracle.basis_eval(b, pts)
    want, wgrad = brute_basis(b, pts)
    assert phi.shape == (24, 200)
    assert np.allclose(phi, want, rtol=1e-13, atol=1e-15)
    assert np.allclose(grad, wgrad, rtol=1e-12, atol=1e-14)
    assert np.all(phi[:, 3] == 0) and np.all(grad[:, :, 3] == 0)


def test_cutoff_is_exact_zero():
    # alpha r^2 >= 50 sets the primitive to exactly 0 (radial1)
    b = basis.Basis(np.zeros((1, 3)), [basis.Shell(0, 0, [2.0], 1, np.zeros((1, 0)))])
    r = np.sqrt(np.array([24.999, 25.0, 25.001]))
    phi, _ = oracle.basis_eval(b, np.column_stack([r, 0 * r, 0 * r]))
    assert phi[0, 0] > 0 and phi[0, 1] == 0 and phi[0, 2] == 0
