"""Screening: the contractions work on gathered runs of 8 basis functions,
leaving out the runs on which the radial cut-off (alpha r^2 >= 50, ParaGauss
modules/orbital_module.f90:1839-1855) makes every basis value exactly zero.
The left-out terms are exact zeros, so screened builds reproduce dense builds
up to the rounding of the regrouped sums (FP64 to 1e-12 relative; the mixed
forms to their own rounding level), and the FP64 oracle (which applies the
same cut-off per point) to the FP64 bar.  The algorithmic work reported by
pgxc_last_stats is checked against n_b counted here from the same cut-off
rule."""
import os

import numpy as np
import pytest

import oracle
from pgxc import XCBuilder, basis, grid, systems
from pgxc.xc import LDA, BF16X3, FP32, FP64, PBE

pytestmark = pytest.mark.gpu

NT = min(16, os.cpu_count() or 1)


def separated_waters(gap_bohr=40.0):
    """Two waters gap_bohr apart: most basis blocks are cut off on most of
    the other molecule's points."""
    e1, x1 = systems.h2o()
    e2, x2 = systems.h2o(center=(gap_bohr, 0.0, 0.0))
    els = list(e1) + list(e2)
    xyz = np.vstack([x1, x2])
    b = basis.seeded_basis(els, xyz, seed=21, shell_set="small")
    g = grid.molecular_grid(els, xyz, nrad=50, nang=302)
    C, occ = systems.random_orbitals(b.nbas, 10, seed=22)
    return els, xyz, b, g, C, occ


@pytest.fixture(scope="module")
def dimer():
    return separated_waters()


def morton_perm(P):
    """The library's point order (Morton curve over the bounding box, 10 bits
    per axis, ties in input order)."""
    lo, hi = P.min(0), P.max(0)
    span = np.where(hi > lo, hi - lo, 1.0)
    q = np.clip(np.floor((P - lo) / span * 1023.0), 0, 1023).astype(np.uint64)
    code = np.zeros(len(P), dtype=np.uint64)
    for bit in range(10):
        for k in range(3):
            code |= ((q[:, k] >> np.uint64(bit)) & np.uint64(1)) << np.uint64(3 * bit + k)
    return np.argsort(code, kind="stable")


def screened_flops(b, g):
    """sum over 128-point tiles of 4 B n_b^2 + 16 B n_b, n_b = functions of
    the shells with alpha_min r^2 < 50 somewhere in a 32-point group of the tile."""
    P = g[morton_perm(g[:, :3]), :3]
    n = len(P)
    f = 0.0
    amin = np.array([s.exps.min() for s in b.shells])
    nfun = np.array([s.nfunc for s in b.shells])
    atom = np.array([s.atom for s in b.shells])
    for t0 in range(0, n, 128):
        T = P[t0:t0 + 128]
        B = len(T)
        sig = np.zeros(len(b.shells), bool)
        for q0 in range(0, B, 32):
            Q = T[q0:q0 + 32]
            r2 = ((Q[:, None, :] - b.atom_xyz[None, :, :]) ** 2).sum(-1).min(0)
            sig |= amin * r2[atom] < 50.0 * (1 + 1e-6)
        nb = nfun[sig].sum()
        f += 4.0 * B * nb * nb + 16.0 * B * nb
    return f


# relative agreement of screened and dense builds: FP64 rounding of regrouped
# sums; FP32 / split-BF16 accumulate each MFMA K step in FP32
REL = {FP64: 1e-12, FP32: 2e-6, BF16X3: 2e-6}


@pytest.mark.parametrize("prec", [FP64, FP32, BF16X3])
def test_screened_equals_dense(dimer, prec):
    _, _, b, g, C, occ = dimer
    out = []
    for screen in (True, False):
        with XCBuilder(b, g, screen=screen) as xc:
            xc.set_orbitals(C, occ)
            out.append(xc.build(PBE, prec))
    (h1, e1, q1), (h2, e2, q2) = out
    print("dE=%.3e dV=%.3e" % (e1 - e2, np.max(np.abs(h1 - h2))), flush=True)
    assert abs(e1 - e2) <= REL[prec] * abs(e2)
    assert np.max(np.abs(h1 - h2)) <= REL[prec] * np.max(np.abs(h2))
    assert abs(q1 - q2) <= REL[prec] * abs(q2)


def test_screened_incremental_equals_dense(dimer):
    _, _, b, g, C, occ = dimer
    C2 = systems.perturbed_orbitals(C, 1e-3, seed=23)
    out = []
    for screen in (True, False):
        with XCBuilder(b, g, screen=screen) as xc:
            xc.set_orbitals(C, occ)
            xc.build(PBE, FP64, fetch=False)
            xc.set_orbitals(C2, occ, incremental=True)
            out.append(xc.build(PBE, BF16X3))
    # the incremental difference is ~1e-3 of the result: its rounding is far smaller
    assert abs(out[0][1] - out[1][1]) <= 1e-10 * abs(out[1][1])
    assert np.max(np.abs(out[0][0] - out[1][0])) <= 1e-10 * np.max(np.abs(out[1][0]))


def test_screened_fp64_matches_oracle(dimer):
    _, _, b, g, C, occ = dimer
    with XCBuilder(b, g) as xc:
        xc.set_orbitals(C, occ)
        ham, e, q = xc.build(PBE, FP64)
    hr, er, qr = oracle.xc_build(b, g, oracle.MASK_PBE, C=C, occ=occ, veclen=128, nthreads=NT)
    print("dE=%.3e dV=%.3e" % (e - er, np.max(np.abs(ham - hr))), flush=True)
    assert abs(e - er) <= 1e-9
    assert np.max(np.abs(ham - hr)) <= 1e-10
    assert abs(q - qr) <= 1e-9


def test_screened_work_counts(dimer):
    _, _, b, g, C, occ = dimer
    n, npts = b.nbas, g.shape[0]
    dense = 4.0 * npts * n * n + 16.0 * npts * n
    expect = screened_flops(b, g)
    for screen, ref in ((True, expect), (False, dense)):
        with XCBuilder(b, g, screen=screen) as xc:
            xc.set_orbitals(C, occ)
            xc.build(PBE, FP64, fetch=False)
            got = xc.stats()["flops"]
        print("screen=%s flops %.6e expected %.6e" % (screen, got, ref), flush=True)
        assert abs(got - ref) <= 2e-3 * ref  # r^2 rounding may flip a shell at the cut-off
    assert expect < 0.75 * dense  # the cut-off removes a real share of the work here


@pytest.mark.parametrize("prec", [FP64, BF16X3])
def test_basis_cache_reuse_is_bitwise(dimer, prec):
    """Builds that reuse the cached basis values equal builds that evaluate
    them, across precision switches and incremental builds."""
    _, _, b, g, C, occ = dimer
    C2 = systems.perturbed_orbitals(C, 1e-3, seed=24)
    out = {}
    for cache in (True, False):
        with XCBuilder(b, g, basis_cache=cache, chunk_points=8192) as xc:
            seq, tb = [], []
            for Ck, p, incr in ((C, FP64, False), (C, prec, False), (C2, prec, True),
                                (C, FP64, False), (C2, prec, True)):
                xc.set_orbitals(Ck, occ, incremental=incr)
                seq.append(xc.build(PBE, p))
                tb.append(xc.timing()["basis"])
            out[cache] = seq
            print("cache=%s basis stage ms per build: %s" % (cache, np.round(tb, 4)), flush=True)
            if cache:  # builds 3-5 read cached values: no basis kernel runs
                assert max(tb[2:]) < 0.5 * tb[0]
    for (h1, e1, q1), (h2, e2, q2) in zip(out[True], out[False]):
        assert np.array_equal(h1, h2) and e1 == e2 and q1 == q2


@pytest.mark.parametrize("prec", [FP64, BF16X3])
def test_lda_then_gga_on_one_context(dimer, prec):
    """An LDA build caches only the basis values: a GGA build after it on the
    same context evaluates the gradients and equals a GGA build on a fresh
    context bit for bit (and a further GGA build reuses the cache)."""
    _, _, b, g, C, occ = dimer
    with XCBuilder(b, g, chunk_points=8192) as xc:
        xc.set_orbitals(C, occ)
        xc.build(LDA, prec)
        r1 = xc.build(PBE, prec)
        r2 = xc.build(PBE, prec)
    with XCBuilder(b, g, chunk_points=8192) as xc:
        xc.set_orbitals(C, occ)
        ref = xc.build(PBE, prec)
    for r in (r1, r2):
        assert np.array_equal(r[0], ref[0]) and r[1] == ref[1] and r[2] == ref[2]


def test_prime_block_count_layout(dimer):
    """A grid of 67 2048-point blocks (a prime above 64) fits one chunk of 34
    point splits of 4096 points (one block of padding) and builds what a
    layout of several chunks builds, to rounding."""
    els, xyz, b, _, C, occ = dimer
    g = grid.molecular_grid(els, xyz, nrad=90, nang=302)[:136000]  # a part of a grid is a grid
    assert g.shape[0] == 136000 and -(-g.shape[0] // 2048) == 67
    out = []
    for chunk in (0, 32768):
        with XCBuilder(b, g, chunk_points=chunk) as xc:
            lay = xc.layout()
            print("chunk_points=%d layout %s" % (chunk, lay), flush=True)
            if chunk == 0:
                assert (lay["chunks"], lay["nsplit"], lay["chunk_points"]) == (1, 34, 34 * 4096)
            else:
                assert lay["chunk_points"] <= chunk
            xc.set_orbitals(C, occ)
            out.append(xc.build(PBE, FP64))
    (h1, e1, q1), (h2, e2, q2) = out
    assert abs(e1 - e2) <= 1e-10 * abs(e1) and abs(q1 - q2) <= 1e-10 * abs(q1)
    assert np.max(np.abs(h1 - h2)) <= 1e-11


@pytest.mark.parametrize("prec", [FP64, BF16X3])
def test_one_lane_and_two_lanes_agree(dimer, prec):
    """Chunks on one stream or alternating between two (pgxc_set_lanes) give
    the same build to the rounding of the per-lane H sums; switching back and
    forth keeps working."""
    _, _, b, g, C, occ = dimer
    with XCBuilder(b, g, chunk_points=8192) as xc:
        assert xc.layout()["chunks"] > 2
        out = []
        for lanes in (2, 1, 2, 1):
            xc.set_lanes(lanes)
            xc.set_orbitals(C, occ)
            out.append(xc.build(PBE, prec))
    (h2, e2, q2), (h1, e1, q1), (h2b, e2b, _), (h1b, e1b, _) = out
    assert np.array_equal(h2, h2b) and e2 == e2b and np.array_equal(h1, h1b) and e1 == e1b
    assert abs(e1 - e2) <= 1e-12 * abs(e1) and abs(q1 - q2) <= 1e-12 * abs(q1)
    assert np.max(np.abs(h1 - h2)) <= 1e-12 * max(1.0, np.max(np.abs(h1)))


def test_state_save_and_load_continue_bitwise(dimer):
    """An incremental build continued from a saved state in a new builder of
    the same geometry (a restarted SCF) gives the bits of the uninterrupted
    one; a builder of another grid refuses the state."""
    from pgxc._lib import PgxcError
    _, _, b, g, C, occ = dimer
    C1 = systems.perturbed_orbitals(C, 1e-3, seed=31)
    C2 = systems.perturbed_orbitals(C1, 1e-3, seed=32)
    with XCBuilder(b, g, chunk_points=8192) as xa:
        xa.set_orbitals(C, occ)
        xa.build(PBE, FP64)
        xa.set_orbitals(C1, occ, incremental=True)
        xa.build(PBE, BF16X3)
        blob = xa.save_state()
        xa.set_orbitals(C2, occ, incremental=True)
        ref = xa.build(PBE, BF16X3)
    with XCBuilder(b, g, chunk_points=8192) as xb:
        xb.load_state(blob)
        xb.set_orbitals(C2, occ, incremental=True)
        out = xb.build(PBE, BF16X3)
    assert np.array_equal(ref[0], out[0]) and ref[1] == out[1] and ref[2] == out[2]
    with XCBuilder(b, g[:-4096], chunk_points=8192) as xc:
        with pytest.raises(PgxcError):
            xc.load_state(blob)


def odd_molecule():
    """Six atoms, 75 functions: the last run of 8 is partly padding."""
    els, xyz = systems.random_molecule(6, seed=1)
    b = basis.seeded_basis(els, xyz, seed=1, shell_set="small")
    g = grid.molecular_grid(els, xyz, nrad=40, nang=194)
    C, occ = systems.random_orbitals(b.nbas, 12, seed=2)
    return b, g, C, occ


@pytest.mark.parametrize("functional", [PBE, LDA])
def test_split_w_from_basis_evaluation_equals_reading(functional, monkeypatch):
    """The split-BF16 W pass evaluates the basis functions again instead of
    reading the stored values and gradients back; the test switch
    PGXC_W_READ=1 makes a context's plans read them.  Incremental split-BF16
    builds on several chunks and both lanes: the same bits either way."""
    els, xyz = systems.water_cluster(4, seed=3)
    b = basis.seeded_basis(els, xyz, seed=4, shell_set="medium")
    g = grid.molecular_grid(els, xyz, nrad=40, nang=302, partition="gpu")
    C, occ = systems.random_orbitals(b.nbas, 20, seed=5)
    C1 = systems.perturbed_orbitals(C, 1e-3, seed=6)
    out = {}
    for read in (False, True):
        if read:
            monkeypatch.setenv("PGXC_W_READ", "1")
        else:
            monkeypatch.delenv("PGXC_W_READ", raising=False)
        with XCBuilder(b, g, chunk_points=32768) as xc:
            assert xc.layout()["chunks"] > 2
            xc.set_orbitals(C, occ)
            xc.build(functional, FP64)
            xc.set_orbitals(C1, occ, incremental=True)
            out[read] = xc.build(functional, BF16X3)
    assert np.array_equal(out[False][0], out[True][0])
    assert out[False][1] == out[True][1] and out[False][2] == out[True][2]


def test_unwritten_basis_blocks_with_wide_point_splits(monkeypatch):
    """A chunk above 256 Ki points makes a point split of the potential span
    more than 32 tiles, past what its per-function tile masks hold: the
    potentials then read every function of their chunks, and the basis and W
    passes write the zeros of the functions outside a tile's list again.
    Buffers poisoned with NaN bytes, all three precisions against the oracle."""
    monkeypatch.setenv("PGXC_POISON_BASIS", "1")
    els, xyz = systems.water_cluster(5, seed=3)  # 306 530 points: 48-tile splits in one chunk
    b = basis.seeded_basis(els, xyz, seed=4, shell_set="medium")
    g = grid.molecular_grid(els, xyz, nrad=75, nang=302)
    C, occ = systems.random_orbitals(b.nbas, 25, seed=5)
    hr, er, qr = oracle.xc_build(b, g, oracle.MASK_PBE, C=C, occ=occ, nthreads=NT)
    with XCBuilder(b, g, chunk_points=524288) as xc:
        lay = xc.layout()
        assert lay["chunks"] == 1 and lay["chunk_points"] // lay["nsplit"] > 32 * 128
        xc.set_orbitals(C, occ)
        for prec, rel in ((FP64, 0.0), (FP32, 2e-6), (BF16X3, 5e-5)):
            ham, e, q = xc.build(PBE, prec)
            dv = np.max(np.abs(ham - hr))
            print("split %d points prec=%d dE=%.2e dV=%.2e"
                  % (lay["chunk_points"] // lay["nsplit"], prec, e - er, dv), flush=True)
            assert np.all(np.isfinite(ham)) and np.isfinite(e) and np.isfinite(q)
            assert abs(e - er) <= 1e-9 + rel * abs(er)
            assert dv <= 1e-10 + 10 * rel * np.max(np.abs(hr))
            assert abs(q - qr) <= 1e-9 + rel * abs(qr)


@pytest.mark.parametrize("system", ["dimer", "odd"])
@pytest.mark.parametrize("cache", [False, True])
def test_unwritten_basis_blocks_are_never_read(dimer, system, cache, monkeypatch):
    """The basis stage writes only the (32-point group x run of 8) blocks a
    later stage reads, the split only those of the values.  With every basis
    buffer filled with NaN bytes first (PGXC_POISON_BASIS), LDA and GGA builds
    in every precision, on several chunks, still match the oracle, and the
    nuclear gradient (whose products read every block, so it evaluates the
    basis densely) stays finite: no stage reads a block left unwritten."""
    monkeypatch.setenv("PGXC_POISON_BASIS", "1")
    if system == "dimer":
        _, _, b, g, C, occ = dimer
    else:
        b, g, C, occ = odd_molecule()
    refs = {fn: oracle.xc_build(b, g, mask, C=C, occ=occ, nthreads=NT)
            for fn, mask in ((LDA, oracle.MASK_LDA), (PBE, oracle.MASK_PBE))}
    with XCBuilder(b, g, basis_cache=cache, chunk_points=8192) as xc:
        assert xc.layout()["chunks"] > 1
        xc.set_orbitals(C, occ)
        for fn in (LDA, PBE):
            hr, er, qr = refs[fn]
            for prec, rel in ((FP64, 0.0), (FP32, 2e-6), (BF16X3, 5e-5)):
                ham, e, q = xc.build(fn, prec)
                dv = np.max(np.abs(ham - hr))
                print("%s nbas=%d cache=%s fn=%#x prec=%d dE=%.2e dV=%.2e"
                      % (system, b.nbas, cache, fn, prec, e - er, dv), flush=True)
                assert np.all(np.isfinite(ham)) and np.isfinite(e) and np.isfinite(q)
                assert abs(e - er) <= 1e-9 + rel * abs(er)
                assert dv <= 1e-10 + 10 * rel * np.max(np.abs(hr))
                assert abs(q - qr) <= 1e-9 + rel * abs(qr)
        xc.build(PBE, FP64)
        gr = xc.gradient()
        assert np.all(np.isfinite(gr))
        again = xc.build(PBE, FP64)  # after the dense pass, the same bits
        assert np.array_equal(again[0], xc.build(PBE, FP64)[0])
