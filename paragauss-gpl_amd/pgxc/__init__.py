"""pgxc -- host-side mirror of ParaGauss's XC grid integration (xc_build ->
calc_xcks_nl) over the MI355X C ABI of include/pgxc.h.

Modules: basis (shells, legacy norms), systems (seeded molecules and
densities), grid (Pople radial x Lebedev x Becke), xc (XCBuilder over the
C ABI), lebedev (302-point rule).
"""
from .basis import Basis, Shell, legacy_norm, seeded_basis  # noqa: F401
from .xc import (BF16X3, FP32, FP64, LDA, PBE, Batch, XCBuilder,  # noqa: F401
                 becke_weights, unpack)
