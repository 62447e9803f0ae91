"""Pins the oracle's functional terms.

PW92 and PBE-c: against the compiled reference modules (oracle/_ref, built
from modules/pw_ldac_module.f90 and modules/pbe_ggcxc_module.f90) where they
are built, and always against the golden vectors tests/golden/functional.npz
that tests/golden/make_golden.py wrote from them.
Slater and PBE-x (modules/exchange.f90, not compilable standalone): closed
forms from the textbook definitions with the reference's constants.
All terms: central finite differences of F give the returned derivatives.
"""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden", "functional.npz")
PI = 3.14159265358979324


def lattice():
    rng = np.random.default_rng(7)
    rho = 10 ** rng.uniform(-14, 3, 600)
    s = rng.uniform(0, 50, 600)
    rho[:4] = [0.0, 1e-31, 1e-25, 1e-20]
    gam = s ** 2 * 4 * (3 * PI ** 2) ** (2 / 3) * rho ** (8 / 3)
    return rho, gam


def rel(a, b, floor):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))


@pytest.mark.skipif(oracle.ref_lib() is None, reason="oracle/_ref not built (no reference tree)")
def test_pw92_matches_reference_build():
    rho, _ = lattice()
    f, fr = oracle.pw92(rho)
    g, gr = oracle.ref_pw92(rho)
    assert rel(f, g, 1e-300) < 1e-13
    assert rel(fr, gr, 1e-300) < 1e-13


@pytest.mark.skipif(oracle.ref_lib() is None, reason="oracle/_ref not built (no reference tree)")
def test_pbec_matches_reference_build():
    rho, gam = lattice()
    f, fr, fg = oracle.pbec(rho, gam)
    g, gr, gg = oracle.ref_pbec(rho, gam)
    # absolute floors: the derivatives vanish like 1/gamma at s -> 50
    assert np.max(np.abs(f - g) / np.maximum(np.abs(g), 1e-20)) < 1e-12
    assert np.max(np.abs(fr - gr) / np.maximum(np.abs(gr), 1e-20)) < 1e-12
    assert np.max(np.abs(fg - gg) / np.maximum(np.abs(gg), 1e-14)) < 1e-9


def test_golden_functional_vectors():
    d = np.load(GOLD)
    f, fr = oracle.pw92(d["rho"])
    assert rel(f, d["pw92_f"], 1e-300) < 1e-13
    assert rel(fr, d["pw92_fr"], 1e-300) < 1e-13
    f, fr, fg = oracle.pbec(d["rho"], d["gam"])
    assert np.max(np.abs(f - d["pbec_f"]) / np.maximum(np.abs(d["pbec_f"]), 1e-20)) < 1e-12
    assert np.max(np.abs(fr - d["pbec_fr"]) / np.maximum(np.abs(d["pbec_fr"]), 1e-20)) < 1e-12
    assert np.max(np.abs(fg - d["pbec_fg"]) / np.maximum(np.abs(d["pbec_fg"]), 1e-14)) < 1e-9


def test_xalpha_closed_form():
    rho, _ = lattice()
    f, fr = oracle.xalpha(rho)
    cx = -0.75 * (3 / PI) ** (1 / 3)
    pos = rho > 0
    assert np.all(f[~pos] == 0) and np.all(fr[~pos] == 0)
    assert rel(f[pos], cx * rho[pos] ** (4 / 3), 1e-300) < 1e-14
    assert rel(fr[pos], cx * 4 / 3 * rho[pos] ** (1 / 3), 1e-300) < 1e-14


def test_pbex_textbook_enhancement():
    # F_x^PBE = e_x^LDA (1 + kappa - kappa / (1 + mu s^2 / kappa)); the
    # reference adds only the excess (the "1" is the Slater term).
    rho, gam = lattice()
    sel = rho >= 1e-20
    f, fr, fg = oracle.pbex(rho, gam)
    kappa, mu = 0.804, 0.066725 * PI ** 2 / 3
    s2 = gam[sel] / (4 * (3 * PI ** 2) ** (2 / 3) * rho[sel] ** (8 / 3))
    ex = -0.75 * (3 / PI) ** (1 / 3) * rho[sel] ** (4 / 3)
    Fx = kappa - kappa / (1 + mu * s2 / kappa)
    assert rel(f[sel], ex * Fx, 1e-300) < 1e-12
    assert np.all(f[~sel] == 0) and np.all(fg[~sel] == 0)


def test_pw92_published_value():
    # PW92 paramagnetic correlation energy per electron at rs = 1 and 2
    # (Perdew & Wang 1992 fit: -0.0598 and -0.0448 hartree to 3 digits)
    for rs, ec in ((1.0, -0.0598), (2.0, -0.0448)):
        n = 3 / (4 * PI * rs ** 3)
        f, _ = oracle.pw92(np.array([n]))
        assert abs(f[0] / n - ec) < 5e-4


@pytest.mark.parametrize("term", ["xalpha", "pw92", "pbex", "pbec", "pbe"])
def test_derivatives_by_central_differences(term):
    rng = np.random.default_rng(3)
    rho = 10 ** rng.uniform(-4, 2, 200)
    s = rng.uniform(0.01, 5, 200)
    gam = s ** 2 * 4 * (3 * PI ** 2) ** (2 / 3) * rho ** (8 / 3)

    def F(r, g):
        if term == "xalpha":
            return oracle.xalpha(r)[0]
        if term == "pw92":
            return oracle.pw92(r)[0]
        if term == "pbex":
            return oracle.pbex(r, g)[0]
        if term == "pbec":
            return oracle.pbec(r, g)[0]
        return oracle.functional(oracle.MASK_PBE, r, g)[0]

    if term in ("xalpha", "pw92"):
        fr = (oracle.xalpha if term == "xalpha" else oracle.pw92)(rho)[1]
        fg = None
    elif term == "pbe":
        _, fr, fg = oracle.functional(oracle.MASK_PBE, rho, gam)
    else:
        _, fr, fg = getattr(oracle, term)(rho, gam)
    h = 1e-5 * rho
    num = (F(rho + h, gam) - F(rho - h, gam)) / (2 * h)
    assert np.max(np.abs(num - fr) / (np.abs(fr) + 1e-12)) < 1e-6
    if fg is not None:
        hg = 1e-5 * gam
        num = (F(rho, gam + hg) - F(rho, gam - hg)) / (2 * hg)
        # in the sum, exchange and correlation gamma-derivatives cancel in part
        tol = 1e-5 if term == "pbe" else 1e-6
        assert np.max(np.abs(num - fg) / (np.abs(fg) + 1e-12)) < tol


def test_functional_is_sum_of_terms():
    rho, gam = lattice()
    f, fr, fg = oracle.functional(oracle.MASK_PBE, rho, gam)
    parts = [oracle.xalpha(rho), oracle.pw92(rho), oracle.pbex(rho, gam), oracle.pbec(rho, gam)]
    assert np.array_equal(f, ((parts[0][0] + parts[1][0]) + parts[2][0]) + parts[3][0])
    assert np.array_equal(fg, parts[2][2] + parts[3][2])


def test_functional_rejects_pbec_without_pw92():
    with pytest.raises(ValueError):
        oracle.functional((1 << 1) | (1 << 18), np.ones(2), np.ones(2))
