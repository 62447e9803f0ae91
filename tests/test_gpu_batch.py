"""GPU: batched builds of several molecules (one launch per stage, hipGraph
replay, cost-ordered schedules) give the same results as one-at-a-time
builds, bit for bit, and match the FP64 oracle."""
import os

import numpy as np
import pytest

import oracle
from pgxc import Batch, XCBuilder, basis, grid, systems
from pgxc.xc import BF16X3, FP32, FP64, PBE

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)


def molecules(n=4, seed=0):
    out = []
    rng = np.random.default_rng(seed)
    for i in range(n):
        na = int(rng.integers(3, 7))
        els, xyz = systems.random_molecule(na, seed=100 + i)
        b = basis.seeded_basis(els, xyz, seed=200 + i, shell_set="small")
        g = grid.molecular_grid(els, xyz, nrad=int(rng.integers(20, 30)), nang=302)
        C, occ = systems.random_orbitals(b.nbas, max(2, b.nbas // 5), seed=300 + i)
        out.append((b, g, C, occ))
    return out


@pytest.fixture(scope="module")
def mols():
    return molecules()


def make_builders(mols, chunk=2048):
    bs = []
    for (b, g, C, occ) in mols:
        xb = XCBuilder(b, g, tile=64, chunk_points=chunk)
        xb.set_orbitals(C, occ)
        bs.append(xb)
    return bs


def test_batch_equals_single_builds_bitwise(mols):
    singles = []
    for (b, g, C, occ) in mols:
        with XCBuilder(b, g, tile=64, chunk_points=2048) as xb:
            xb.set_orbitals(C, occ)
            singles.append(xb.build(PBE, FP64))
    bs = make_builders(mols)
    with Batch(bs) as batch:
        batch.build(PBE, FP64)
        res = batch.results()
    for r, s in zip(res, singles):
        assert np.array_equal(r[0], s[0]) and r[1] == s[1] and r[2] == s[2]
    for xb in bs:
        xb.close()


def test_batch_matches_oracle(mols):
    bs = make_builders(mols)
    with Batch(bs) as batch:
        batch.build(PBE, FP64)
        res = batch.results()
    for (b, g, C, occ), (ham, e, q) in zip(mols, res):
        hr, er, qr = oracle.xc_build(b, g, oracle.MASK_PBE, C=C, occ=occ, nthreads=NT)
        print("dE=%.2e dV=%.2e" % (e - er, np.max(np.abs(ham - hr))), flush=True)
        assert abs(e - er) <= 1e-9 and np.max(np.abs(ham - hr)) <= 1e-10 and abs(q - qr) <= 1e-9
    for xb in bs:
        xb.close()


@pytest.mark.parametrize("prec", [FP64, FP32, BF16X3])
def test_graph_replay_and_schedule_do_not_change_bits(mols, prec):
    bs = make_builders(mols)
    with Batch(bs) as batch:
        batch.build(PBE, prec)
        ref = batch.results()
        rng = np.random.default_rng(5)
        for rep in range(3):  # several replays with eager work in between
            batch.set_priority(rng.uniform(size=len(bs)) if rep else None)
            batch.build(PBE, prec, graph=True)
            got = batch.results()
            for r, s in zip(got, ref):
                assert np.array_equal(r[0], s[0]) and r[1] == s[1] and r[2] == s[2]
            bs[0].build(PBE, prec)  # eager build on one member between replays
    for xb in bs:
        xb.close()


def test_batch_with_mixed_lane_counts(mols):
    """One multi-chunk member set to one lane, another left at two: the batch
    then runs one lane for all, each member's chunks on the buffers of the
    stream that runs them, and every member gets the bits of its own build
    at one lane (no two chunks of a member share buffers concurrently)."""
    singles = []
    for (b, g, C, occ) in mols:
        with XCBuilder(b, g, tile=64, chunk_points=2048) as xb:
            assert xb.layout()["chunks"] > 1
            xb.set_lanes(1)
            xb.set_orbitals(C, occ)
            singles.append(xb.build(PBE, FP64))
    bs = make_builders(mols)
    bs[0].set_lanes(1)  # the others keep two lanes
    with Batch(bs) as batch:
        for graph in (False, True, True):
            batch.build(PBE, FP64, graph=graph)
            for r, s in zip(batch.results(), singles):
                assert np.array_equal(r[0], s[0]) and r[1] == s[1] and r[2] == s[2]
    for xb in bs:
        xb.close()
