"""Writes the golden vectors of tests/golden/ from the compiled reference
modules (oracle/_ref/libpgref.so, built by `make -f oracle/ref.mk` from the
unmodified ParaGauss sources).  Run from the repository root:

    python tests/golden/make_golden.py

functional.npz:  PW92 (pw_ldac, iop = 1) and PBE-c (pbe_ggcc "pbe", ispin = 1)
                 on a (rho, s) lattice covering rho in [1e-14, 1e3], s in [0, 50]
                 and the rho = 0 / rho < 1e-20 / rho < 1e-30 special cases.
harmonics.npz:   solid_harmonics_calculate + _grads for l <= 4 on seeded points.
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
    rng = np.random.default_rng(2024)
    rho = 10 ** rng.uniform(-14, 3, 400)
    s = rng.uniform(0, 50, 400)
    rho[:5] = [0.0, 1e-31, 1e-25, 1e-20, 5e-21]
    gam = s ** 2 * 4 * (3 * PI ** 2) ** (2 / 3) * rho ** (8 / 3)
    f, fr = oracle.ref_pw92(rho)
    g, gr, gg = oracle.ref_pbec(rho, gam)
    np.savez_compressed(os.path.join(HERE, "functional.npz"), rho=rho, gam=gam, pw92_f=f,
                        pw92_fr=fr, pbec_f=g, pbec_fr=gr, pbec_fg=gg)
    pts = rng.normal(size=(64, 3)) * 1.5
    ctr = np.array([0.25, -0.5, 0.75])
    C, G = oracle.ref_solid_harmonics(4, pts, ctr)
    np.savez_compressed(os.path.join(HERE, "harmonics.npz"), pts=pts, center=ctr, C=C, G=G)
    print("wrote functional.npz, harmonics.npz")


if __name__ == "__main__":
    main()
