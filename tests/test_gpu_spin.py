"""Open-shell builds (ispin = 2) on the GPU against the FP64 oracle's
open-shell build on identical inputs (modules/xc_hamiltonian.f90:1428-1460,
:1555-1567): the same bars as the closed-shell parity tests,
|dE_xc| <= 1e-9 Eh and max |dV_xc| <= 1e-10 for FP64, over both spins of
the interleaved output."""
import os

import numpy as np
import pytest

import oracle
from pgxc import XCBuilder, basis, grid, systems
from pgxc.xc import BF16X3, FP32, FP64, LDA, PBE

pytestmark = pytest.mark.gpu

NT = min(16, os.cpu_count() or 1)


def spin_orbitals(nbas, seed):
    """Alpha: 5 orbitals, beta: a rotated set with fractional occupations, so
    the spin densities differ at every point."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(nbas, 6)))
    C = np.stack([Q[:, :5], Q[:, 1:6]], axis=2)
    occ = np.stack([np.ones(5), np.array([1.0, 1.0, 0.8, 0.5, 0.0])], axis=1)
    return C, occ


@pytest.fixture(scope="module")
def case():
    els, xyz = systems.h2o()
    b = basis.seeded_basis(els, xyz, seed=1, shell_set="small")
    g = grid.molecular_grid(els, xyz, nrad=75, nang=302)
    C, occ = spin_orbitals(b.nbas, 4)
    return b, g, C, occ


_REF = {}


def ref(case, mask):
    b, g, C, occ = case
    if mask not in _REF:
        _REF[mask] = oracle.xc_build(b, g, mask, C=C, occ=occ, ispin=2, nthreads=NT)
    return _REF[mask]


def check(res, r, e_tol=1e-9, v_tol=1e-10, q_tol=1e-9):
    ham, e, q = res
    hr, er, qr = r
    assert ham.shape == hr.shape
    dva = np.max(np.abs(ham[0::2] - hr[0::2]))
    dvb = np.max(np.abs(ham[1::2] - hr[1::2]))
    print("dE=%.3e dV_alpha=%.3e dV_beta=%.3e dQ=%.3e" % (e - er, dva, dvb, q - qr), flush=True)
    assert abs(e - er) <= e_tol
    assert dva <= v_tol and dvb <= v_tol
    assert abs(q - qr) <= q_tol


@pytest.mark.parametrize("functional,mask", [(PBE, oracle.MASK_PBE), (LDA, oracle.MASK_LDA)],
                         ids=["PBE", "LDA"])
def test_open_shell_fp64_matches_oracle(case, functional, mask):
    b, g, C, occ = case
    with XCBuilder(b, g) as xc:
        xc.set_orbitals(C, occ)
        res = xc.build(functional, FP64)
    check(res, ref(case, mask))


def test_open_shell_chunked_two_lanes(case):
    b, g, C, occ = case
    with XCBuilder(b, g, chunk_points=16384) as xc:
        assert xc.layout()["chunks"] >= 3
        xc.set_orbitals(C, occ)
        res = xc.build(PBE, FP64)
    check(res, ref(case, oracle.MASK_PBE))


def test_equal_spins_give_the_closed_shell_build(case):
    b, g, C, occ = case
    P = (C[:, :, 0] * occ[:, 0]) @ C[:, :, 0].T
    with XCBuilder(b, g) as xc:
        xc.set_density(P)
        h1, e1, q1 = xc.build(PBE, FP64)
        xc.set_density(np.stack([P / 2, P / 2], axis=2))
        h2, e2, q2 = xc.build(PBE, FP64)
    scale = np.max(np.abs(h1))
    assert abs(e2 - e1) <= 1e-11 * abs(e1)
    assert abs(q2 - q1) <= 1e-12 * abs(q1)
    assert np.max(np.abs(h2[0::2] - h1)) <= 1e-11 * scale
    assert np.max(np.abs(h2[1::2] - h1)) <= 1e-11 * scale


def test_open_shell_split_bf16(case):
    # the closed-shell mixed-precision bar: relative 5e-5 of |E_xc| and the
    # charge, 10x that of max |V| (this basis reaches exponents of 5000)
    b, g, C, occ = case
    r = ref(case, oracle.MASK_PBE)
    with XCBuilder(b, g) as xc:
        xc.set_orbitals(C, occ)
        res = xc.build(PBE, BF16X3)
    rel = 5e-5
    check(res, r, e_tol=rel * abs(r[1]), v_tol=10 * rel * np.max(np.abs(r[0])),
          q_tol=rel * abs(r[2]))


def test_open_shell_split_w_from_basis_evaluation_equals_reading(case, monkeypatch):
    """Both spins' split-BF16 W from one evaluation of the basis (the default)
    or from the stored values read back (PGXC_W_READ=1): the same bits."""
    b, g, C, occ = case
    out = {}
    for read in (False, True):
        if read:
            monkeypatch.setenv("PGXC_W_READ", "1")
        else:
            monkeypatch.delenv("PGXC_W_READ", raising=False)
        with XCBuilder(b, g) as xc:
            xc.set_orbitals(C, occ)
            out[read] = xc.build(PBE, BF16X3)
    assert np.array_equal(out[False][0], out[True][0])
    assert out[False][1] == out[True][1] and out[False][2] == out[True][2]


def test_open_shell_incremental_and_restart(case):
    """A full build, then an incremental one with changed spin densities:
    the FP64 bar against the oracle's full build of the new densities; a
    state saved after the first build and loaded into a fresh builder
    continues bit for bit."""
    # perturbed orbitals keep both spin densities non-negative (a negative
    # spin density has no real spin-polarised correlation energy)
    b, g, C, occ = case
    C1 = np.stack([systems.perturbed_orbitals(C[:, :, s], 1e-2, seed=9 + s) for s in range(2)],
                  axis=2)
    with XCBuilder(b, g) as xc:
        xc.set_orbitals(C, occ)
        xc.build(PBE, FP64)
        state = xc.save_state()
        xc.set_orbitals(C1, occ, incremental=True)
        inc = xc.build(PBE, FP64)
    with XCBuilder(b, g) as xc2:
        xc2.load_state(state)
        assert xc2.ispin == 2
        xc2.set_orbitals(C1, occ, incremental=True)
        inc2 = xc2.build(PBE, FP64)
    check(inc, oracle.xc_build(b, g, oracle.MASK_PBE, C=C1, occ=occ, ispin=2, nthreads=NT))
    assert inc2[1] == inc[1] and np.array_equal(inc2[0], inc[0])


def test_incremental_with_other_spin_count_is_refused(case):
    b, g, C, occ = case
    with XCBuilder(b, g) as xc:
        xc.set_orbitals(C, occ)
        xc.build(PBE, FP64)
        P = (C[:, :, 0] * occ[:, 0]) @ C[:, :, 0].T
        with pytest.raises(Exception):
            xc.set_density(P, incremental=True)


def test_batch_of_open_and_closed_shells_is_bitwise(case):
    """One batch holding an open-shell and a closed-shell molecule gives each
    the bits of its own single build."""
    from pgxc import Batch
    b, g, C, occ = case
    els, xyz = systems.random_molecule(4, seed=103)
    b2 = basis.seeded_basis(els, xyz, seed=203, shell_set="small")
    g2 = grid.molecular_grid(els, xyz, nrad=24, nang=302)
    C2, occ2 = systems.random_orbitals(b2.nbas, 4, seed=303)
    jobs = [(b, g, (C, occ)), (b2, g2, (C2, occ2))]
    singles = []
    for (bb, gg, (cc, oo)) in jobs:
        with XCBuilder(bb, gg, tile=64, chunk_points=8192) as xb:
            xb.set_orbitals(cc, oo)
            singles.append(xb.build(PBE, BF16X3))
    bs = []
    for (bb, gg, (cc, oo)) in jobs:
        xb = XCBuilder(bb, gg, tile=64, chunk_points=8192)
        xb.set_orbitals(cc, oo)
        bs.append(xb)
    with Batch(bs) as batch:
        batch.build(PBE, BF16X3)
        res = batch.results()
    for r, s in zip(res, singles):
        assert np.array_equal(r[0], s[0]) and r[1] == s[1] and r[2] == s[2]
    for xb in bs:
        xb.close()


def test_fortran_driver_open_shell(case, tmp_path):
    """The Fortran caller with eigvec_occ(nbas, nocc, 2), occ_num_occ(nocc, 2)
    and an interleaved ham_xc_arr of 2 nbas (nbas + 1) / 2 entries."""
    import subprocess
    b, g, C, occ = case
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([b.n_atoms, b.n_shells, b.nbas, C.shape[1], 128, FP64, PBE, 2],
                 dtype=np.int32).tofile(f)
        np.array([g.shape[0]], dtype=np.int64).tofile(f)
        np.ascontiguousarray(b.atom_xyz).tofile(f)
        np.ascontiguousarray(b.shell_info, dtype=np.int32).tofile(f)
        np.array([len(b.exps), len(b.contr)], dtype=np.int32).tofile(f)
        b.exps.tofile(f)
        b.norms.tofile(f)
        np.ascontiguousarray(b.contr).tofile(f)
        np.asarray(g, dtype=np.float64).ravel(order="F").tofile(f)
        np.asarray(C, dtype=np.float64).ravel(order="F").tofile(f)
        np.asarray(occ, dtype=np.float64).ravel(order="F").tofile(f)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "paragauss-gpl_amd", "fortran", "pgxc_driver")
    out = subprocess.run([exe, str(fin), str(fout)], capture_output=True, text=True, timeout=300)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    raw = np.fromfile(fout, dtype=np.float64)
    nh = b.nbas * (b.nbas + 1)  # two spins
    check((raw[2:2 + nh], raw[0], raw[1]), ref(case, oracle.MASK_PBE))
    assert abs(raw[2 + nh:6 + nh].sum() - raw[0]) <= 1e-12 * abs(raw[0])
    # after the restart (state words, new setup, incremental build with
    # dP = 0): the same bits as the build before it
    tail = raw[7 + nh + 3 * b.n_atoms:]
    assert tail.size == 2 + nh
    assert tail[0] == raw[0] and tail[1] == raw[1] and np.array_equal(tail[2:], raw[2:2 + nh])


def test_fully_polarised_and_empty_spin(case):
    """Beta spin empty (a fully polarised density, as for a doublet radical
    with every electron in alpha): zeta = 1 everywhere, the DEPS-regularised
    branches of PW92 and PBE-c; and a build with no electrons at all."""
    b, g, C, occ = case
    occ2 = occ.copy()
    occ2[:, 1] = 0.0
    r = oracle.xc_build(b, g, oracle.MASK_PBE, C=C, occ=occ2, ispin=2, nthreads=NT)
    with XCBuilder(b, g) as xc:
        xc.set_orbitals(C, occ2)
        res = xc.build(PBE, FP64)
        assert np.all(np.isfinite(res[0]))
        check(res, r)
        xc.set_orbitals(C[:, :0, :], occ[:0, :])
        ham, e, q = xc.build(PBE, FP64)
    # PW92 clamps n below DEPS = 1e-30 to DEPS: a vanishing, not zero, energy
    assert abs(e) <= 1e-20 and q == 0.0 and not np.any(ham)


def test_open_shell_water_cluster_medium_basis():
    """(H2O)4 in the 33-function water basis, 40 x 302 grid in chunks of 16 Ki
    points on two lanes: screening, several chunks and both spins' units in
    every stage; FP64 against the oracle's open-shell build."""
    els, xyz = systems.water_cluster(4, seed=3)
    b = basis.seeded_basis(els, xyz, seed=4, shell_set="medium")
    g = grid.molecular_grid(els, xyz, nrad=40, nang=302, partition="gpu")
    rng = np.random.default_rng(6)
    Q, _ = np.linalg.qr(rng.normal(size=(b.nbas, 21)))
    C = np.stack([Q[:, :20], Q[:, 1:21]], axis=2)
    occ = np.stack([np.ones(20), np.r_[np.ones(19), 0.0]], axis=1)
    with XCBuilder(b, g, chunk_points=16384) as xc:
        assert xc.layout()["chunks"] >= 3
        xc.set_orbitals(C, occ)
        res = xc.build(PBE, FP64)
    check(res, oracle.xc_build(b, g, oracle.MASK_PBE, C=C, occ=occ, ispin=2, nthreads=NT))


@pytest.mark.parametrize("prec", [FP32, BF16X3], ids=["FP32", "BF16X3"])
@pytest.mark.parametrize("norm", [1e-2, 1e-4])
def test_open_shell_incremental_mixed_error_tracks_dP(case, norm, prec):
    """Incremental mixed-precision open-shell builds from an FP64 state: the
    closed-shell bar, the FP64 floor plus a multiple of ||dP||."""
    b, g, C, occ = case
    C1 = np.stack([systems.perturbed_orbitals(C[:, :, s], norm, seed=30 + s) for s in range(2)],
                  axis=2)
    with XCBuilder(b, g) as xc:
        xc.set_orbitals(C, occ)
        xc.build(PBE, FP64)
        xc.set_orbitals(C1, occ, incremental=True)
        inc = xc.build(PBE, prec)
    r = oracle.xc_build(b, g, oracle.MASK_PBE, C=C1, occ=occ, ispin=2, nthreads=NT)
    k = 1e-3 if prec == FP32 else 5e-3
    check(inc, r, e_tol=1e-9 + k * norm, v_tol=1e-10 + k * norm, q_tol=1e-9 + k * norm)
