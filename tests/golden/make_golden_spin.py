"""Writes tests/golden/functional_spin.npz from the compiled reference modules
(oracle/_ref/libpgref.so, built by `make -f oracle/ref.mk` from the
unmodified ParaGauss sources).  Run from the repository root:

    python tests/golden/make_golden_spin.py

PW92 (pw_ldac, iop = PWLDAC_UNRESTR) and PBE-c (pbe_ggcc "pbe", ispin = 2) on
seeded spin densities n_a, n_b in [1e-8, 1e3] with spin polarisations from
fully unpolarised to nearly fully polarised, reduced gradients s in [0, 20]
and gab between -sqrt(gaa gbb) and sqrt(gaa gbb), plus points with one spin
density zero.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

PI = 3.14159265358979324


def main():
    if oracle.ref_lib() is None:
        raise SystemExit("oracle/_ref/libpgref.so missing: make -f oracle/ref.mk")
    rng = np.random.default_rng(2025)
    n = 400
    nt = 10 ** rng.uniform(-8, 3, n)
    zeta = rng.uniform(-1, 1, n)
    zeta[:40] = 0.0
    zeta[40:60] = rng.choice([-1.0, 1.0], 20) * (1 - 10 ** rng.uniform(-9, -3, 20))
    na, nb = 0.5 * nt * (1 + zeta), 0.5 * nt * (1 - zeta)
    na[60:70] = 0.0
    nb[70:80] = 0.0
    k = 4 * (3 * PI ** 2) ** (2 / 3)
    gaa = rng.uniform(0, 20, n) ** 2 * k * (2 * na) ** (8 / 3) / 4
    gbb = rng.uniform(0, 20, n) ** 2 * k * (2 * nb) ** (8 / 3) / 4
    gab = rng.uniform(-1, 1, n) * np.sqrt(gaa * gbb)
    pf, pa, pb = oracle.ref_pw92_spin(na, nb)
    cf, ca, cb, cgaa, cgbb, cgab = oracle.ref_pbec_spin(na, nb, gaa, gbb, gab)
    np.savez_compressed(os.path.join(HERE, "functional_spin.npz"), na=na, nb=nb, gaa=gaa,
                        gbb=gbb, gab=gab, pw92_f=pf, pw92_fa=pa, pw92_fb=pb, pbec_f=cf,
                        pbec_fa=ca, pbec_fb=cb, pbec_fgaa=cgaa, pbec_fgbb=cgbb, pbec_fgab=cgab)
    print("wrote functional_spin.npz")


if __name__ == "__main__":
    main()
