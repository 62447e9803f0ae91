"""Spin-polarised functional of the oracle (ispin = 2, or_functional_spin):
pinned against the reference modules' own unrestricted branches (golden
vectors of pw_ldac iop = PWLDAC_UNRESTR and pbe_ggcc ispin = 2, written by
tests/golden/make_golden_spin.py from the compiled reference), against the
closed-shell oracle at zero polarisation, against the spin-scaling identity
of exchange, and against central differences."""
import os

import numpy as np

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden", "functional_spin.npz")


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


def test_golden_spin_vectors():
    d = np.load(GOLD)
    na, nb, gaa, gbb, gab = d["na"], d["nb"], d["gaa"], d["gbb"], d["gab"]
    # points where one spin density vanishes (or is below 1e-9 of the other):
    # there the derivative in that density carries the DEPS-regularised
    # (1 -+ zeta)^(1/3) term, where the two forward-mode evaluations differ at
    # 1e-5 relative; everywhere else the bar is 1e-10 (the rounding of the
    # cancellations near full polarisation and in tiny gradient derivatives)
    edge = (np.minimum(na, nb) <= 1e-9 * np.maximum(na, nb))
    f, fa, fb = oracle.pw92_spin(na, nb)
    for got, ref in ((f, d["pw92_f"]), (fa, d["pw92_fa"]), (fb, d["pw92_fb"])):
        assert np.all(_rel(got, ref)[~edge] <= 1e-10)
        assert np.all(_rel(got, ref)[edge] <= 1e-5)
    cf, ca, cb, cg2 = oracle.pbec_spin(na, nb, gaa, gbb, gab)
    for got, ref in ((cf, d["pbec_f"]), (ca, d["pbec_fa"]), (cb, d["pbec_fb"]),
                     (cg2, d["pbec_fgaa"]), (cg2, d["pbec_fgbb"]), (2 * cg2, d["pbec_fgab"])):
        assert np.all(_rel(got, ref)[~edge] <= 1e-10)
        assert np.all(_rel(got, ref)[edge] <= 1e-5)


def test_unpolarised_limit_is_closed_shell():
    """n_a = n_b = n/2, g_aa = g_bb = g_ab = g/4: f and dF/dn_a equal the
    closed-shell f and dF/dn, and d/dg F(g/4, g/4, g/4) = (dF/dgaa + dF/dgbb +
    dF/dgab)/4 = dF/dg."""
    rng = np.random.default_rng(5)
    n = 10 ** rng.uniform(-6, 2, 300)
    g = rng.uniform(0, 10, 300) ** 2 * n ** (8 / 3)
    for mask in (oracle.MASK_LDA, oracle.MASK_PBE):
        f1, fr1, fg1 = oracle.functional(mask, n, g)
        f2, fr2, fg2 = oracle.functional_spin(mask, n / 2, n / 2, g / 4, g / 4, g / 4)
        assert np.all(_rel(f2, f1) <= 1e-13)
        assert np.all(_rel(fr2[:, 0], fr1) <= 1e-12) and np.all(_rel(fr2[:, 1], fr1) <= 1e-12)
        if mask == oracle.MASK_PBE:
            # exchange and correlation gradient terms cancel where g is
            # small, so the sum carries a few 1e-11 of relative rounding
            tot = (fg2[:, 0] + fg2[:, 1] + fg2[:, 2]) / 4
            assert np.all(_rel(tot, fg1) <= 1e-10)


def test_exchange_spin_scaling():
    """E_x[n_a, n_b] = (E_x[2 n_a] + E_x[2 n_b]) / 2 with g_ss scaled by 4 (the
    closed-shell exchange is pinned by its closed form)."""
    rng = np.random.default_rng(6)
    na, nb = 10 ** rng.uniform(-5, 2, 200), 10 ** rng.uniform(-5, 2, 200)
    gaa, gbb = rng.uniform(0, 5, 200) ** 2 * na ** (8 / 3), rng.uniform(0, 5, 200) ** 2 * nb ** (8 / 3)
    mask = (1 << oracle.XC_XALPHA) | (1 << oracle.XC_PBEX)
    f, fr, fg = oracle.functional_spin(mask, na, nb, gaa, gbb, np.zeros(200))
    fa, fra, fga = oracle.functional(mask, 2 * na, 4 * gaa)
    fb, frb, fgb = oracle.functional(mask, 2 * nb, 4 * gbb)
    assert np.all(_rel(f, 0.5 * (fa + fb)) <= 1e-13)
    assert np.all(_rel(fr[:, 0], fra) <= 1e-13) and np.all(_rel(fr[:, 1], frb) <= 1e-13)
    assert np.all(_rel(fg[:, 0], 2 * fga) <= 1e-13) and np.all(_rel(fg[:, 1], 2 * fgb) <= 1e-13)
    assert np.all(fg[:, 2] == 0.0)


def test_spin_derivatives_by_central_differences():
    rng = np.random.default_rng(7)
    m = 100
    na, nb = 10 ** rng.uniform(-3, 1, m), 10 ** rng.uniform(-3, 1, m)
    gaa, gbb = rng.uniform(0.1, 3, m) ** 2 * na ** (8 / 3), rng.uniform(0.1, 3, m) ** 2 * nb ** (8 / 3)
    gab = rng.uniform(-0.9, 0.9, m) * np.sqrt(gaa * gbb)
    x = [na, nb, gaa, gbb, gab]
    f, fr, fg = oracle.functional_spin(oracle.MASK_PBE, *x)
    an = [fr[:, 0], fr[:, 1], fg[:, 0], fg[:, 1], fg[:, 2]]
    for k in range(5):
        h = 1e-5 * x[k]
        xp = [v.copy() for v in x]
        xm = [v.copy() for v in x]
        xp[k] += h
        xm[k] -= h
        fd = (oracle.functional_spin(oracle.MASK_PBE, *xp)[0]
              - oracle.functional_spin(oracle.MASK_PBE, *xm)[0]) / (2 * h)
        assert np.all(np.abs(fd - an[k]) <= 1e-6 * np.abs(an[k]) + 1e-9 * np.abs(f / x[k])), k
