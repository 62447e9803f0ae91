"""Post-SCF XC energies per functional term (pgxc_energy,
modules/post_scf_module.f90:2204-2420) on the GPU: each term's energy
against the FP64 oracle's build with that term alone (PBEc through PWLDAc +
PBEc minus PWLDAc, the oracle refusing PBEc without its LDA part), their sum
against the full build's E_xc, and the incremental state left untouched."""
import os

import numpy as np
import pytest

import oracle
from pgxc import XCBuilder, basis, grid, systems
from pgxc.xc import BF16X3, FP64, LDA, PBE

pytestmark = pytest.mark.gpu

NT = min(16, os.cpu_count() or 1)
XA, PW, PX, PC = (1 << oracle.XC_XALPHA, 1 << oracle.XC_PWLDAC, 1 << oracle.XC_PBEX,
                  1 << oracle.XC_PBEC)


@pytest.fixture(scope="module")
def case():
    els, xyz = systems.h2o()
    b = basis.seeded_basis(els, xyz, seed=1, shell_set="small")
    g = grid.molecular_grid(els, xyz, nrad=75, nang=302)
    C, occ = systems.random_orbitals(b.nbas, 5, seed=2)
    return b, g, C, occ


def oracle_terms(b, g, **dens):
    e = {m: oracle.xc_build(b, g, m, nthreads=NT, **dens)[1] for m in (XA, PW, PX, PW | PC)}
    q = oracle.xc_build(b, g, XA, nthreads=NT, **dens)[2]
    return {"xalpha": e[XA], "pwldac": e[PW], "pbex": e[PX], "pbec": e[PW | PC] - e[PW]}, q


def test_energy_terms_match_oracle(case):
    b, g, C, occ = case
    ref, qr = oracle_terms(b, g, C=C, occ=occ)
    with XCBuilder(b, g) as xc:
        xc.set_orbitals(C, occ)
        terms, q = xc.energy(PBE, FP64)
        lda, _ = xc.energy(LDA, FP64)
    for k in ref:
        print(k, terms[k], terms[k] - ref[k], flush=True)
        assert abs(terms[k] - ref[k]) <= 1e-9
    assert abs(q - qr) <= 1e-9
    assert lda["xalpha"] == terms["xalpha"] and lda["pwldac"] == terms["pwldac"]
    assert lda["pbex"] == 0.0 and lda["pbec"] == 0.0


def test_energy_terms_sum_to_build_and_leave_state(case):
    b, g, C, occ = case
    C1 = systems.perturbed_orbitals(C, 1e-3, seed=5)
    with XCBuilder(b, g) as xc:
        xc.set_orbitals(C, occ)
        ham, e, q = xc.build(PBE, FP64)
        terms, qe = xc.energy(PBE, FP64)
        assert abs(sum(terms.values()) - e) <= 1e-12 * abs(e)
        assert abs(qe - q) <= 1e-12 * abs(q)
        assert np.array_equal(xc.result()[0], ham) and xc.result()[1] == e
        xc.set_orbitals(C1, occ, incremental=True)
        xc.energy(PBE, BF16X3)
        inc = xc.build(PBE, FP64)
    with XCBuilder(b, g) as xc:
        xc.set_orbitals(C, occ)
        xc.build(PBE, FP64)
        xc.set_orbitals(C1, occ, incremental=True)
        inc_ref = xc.build(PBE, FP64)
    assert np.array_equal(inc[0], inc_ref[0]) and inc[1] == inc_ref[1]


def test_energy_terms_open_shell(case):
    b, g, _, _ = case
    rng = np.random.default_rng(4)
    Q, _ = np.linalg.qr(rng.normal(size=(b.nbas, 6)))
    C = np.stack([Q[:, :5], Q[:, 1:6]], axis=2)
    occ = np.stack([np.ones(5), np.array([1.0, 1.0, 0.8, 0.5, 0.0])], axis=1)
    ref, qr = oracle_terms(b, g, C=C, occ=occ, ispin=2)
    with XCBuilder(b, g) as xc:
        xc.set_orbitals(C, occ)
        terms, q = xc.energy(PBE, FP64)
        mixed, _ = xc.energy(PBE, BF16X3)
    for k in ref:
        print(k, terms[k], terms[k] - ref[k], mixed[k] - ref[k], flush=True)
        assert abs(terms[k] - ref[k]) <= 1e-9
        assert abs(mixed[k] - ref[k]) <= 5e-5 * abs(ref[k]) + 1e-9
    assert abs(q - qr) <= 1e-9
