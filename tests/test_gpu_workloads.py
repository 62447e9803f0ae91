"""GPU parity at the measured sizes: the workloads bench.py times, built
through the C ABI and checked against the FP64 oracle on the host's cores.

  (H2O)32, 1056 functions, ~1.96 M points, 160 occupied:
    grid (Slater radii, int(NRAD R) shells, tabulated Lebedev, Becke on the
    GPU) against the oracle's mkgrid; the FP64 build against the oracle's
    build: |dE_xc| <= 1e-9 Eh, max |dV_xc| <= 1e-10, |dcharge| <= 1e-9;
    FP32 and split-BF16 incremental builds at ||dP||_F = 1e-2, 1e-4, 1e-6
    against the oracle's build of the new density: the FP64 bar plus
    k ||dP||_F (k = 1e-3 FP32, 5e-3 split-BF16), and |dE_xc| <= 1e-6 Eh.
  64 molecules of 3-60 atoms in one batch: graph-replayed FP64 batch builds
    bit for bit equal to one build per molecule; the smallest, the largest
    and a middle molecule against the oracle (the FP64 bar); split-BF16
    incremental batch builds on a new density (||dP||_F = 1e-4) within
    1e-6 Eh of the FP64 builds.
  (H2O)128 + 40-atom solute, 4982 functions, ~8 M points, 640 occupied:
    static, analytic and cost-model schedules bit-identical (FP64 and
    split-BF16); FP64 builds of three slices of the grid against the oracle
    on the same slices (the FP64 bar).
"""
import os

import numpy as np
import pytest

import oracle
from pgxc import Batch, XCBuilder, _lib, grid, systems, workloads
from pgxc.costmodel import default_model
from pgxc.xc import BF16X3, FP32, FP64, PBE

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)


def say(*a):
    print(*a, flush=True)


def check(res, ref, e_tol=1e-9, v_tol=1e-10, q_tol=1e-9, what=""):
    ham, e, q = res
    hr, er, qr = ref
    dv = float(np.max(np.abs(ham - hr)))
    say("%s dE=%.3e max|dV|=%.3e dQ=%.3e (E_xc %.9f)" % (what, e - er, dv, q - qr, er))
    assert abs(e - er) <= e_tol
    assert dv <= v_tol
    assert abs(q - qr) <= q_tol


# ---------------------------------------------------------------- (H2O)32
@pytest.fixture(scope="module")
def h2o32():
    say("(H2O)32: building the workload (grid Becke-partitioned on the GPU)")
    return workloads.h2o32()


_REF = {}


def h2o32_oracle(key, b, g, C, occ):
    if key not in _REF:
        say("(H2O)32 oracle build %s on %d threads" % (key, NT))
        _REF[key] = oracle.xc_build(b, g, oracle.MASK_PBE, C=C, occ=occ, veclen=128, nthreads=NT)
    return _REF[key]


def test_h2o32_grid_matches_oracle(h2o32):
    els, b, g, C, occ = h2o32
    go = oracle.molecular_grid(els, b.atom_xyz, 75, 302, nthreads=NT)
    say("(H2O)32 grid: %d points (oracle %d)" % (g.shape[0], go.shape[0]))
    assert g.shape == go.shape
    assert np.array_equal(g[:, :3], go[:, :3])
    # the partition factor P_owner / sum P lies in [0, 1]: bound its error
    # absolutely, as a fraction of each point's one-centre weight
    w_at = grid.molecular_grid(els, b.atom_xyz, nrad=75, nang=302, partition="none")[:, 3]
    dp = np.abs(g[:, 3] - go[:, 3]) / w_at
    say("max |d partition factor| = %.3e" % dp.max())
    assert dp.max() <= 1e-13


def test_h2o32_fp64_matches_oracle(h2o32):
    els, b, g, C, occ = h2o32
    say("(H2O)32 FP64 GPU build: N=%d, %d points" % (b.nbas, g.shape[0]))
    with XCBuilder(b, g) as xc:
        xc.set_orbitals(C, occ)
        res = xc.build(PBE, FP64)
    check(res, h2o32_oracle("C", b, g, C, occ), what="(H2O)32 FP64")


@pytest.mark.parametrize("norm", [1e-2, 1e-4, 1e-6])
def test_h2o32_incremental_mixed_matches_oracle(h2o32, norm):
    els, b, g, C, occ = h2o32
    C1 = systems.perturbed_orbitals(C, norm, seed=100)
    ref = h2o32_oracle(("C1", norm), b, g, C1, occ)
    with XCBuilder(b, g) as xc:
        for prec, k in ((FP32, 1e-3), (BF16X3, 5e-3)):
            say("(H2O)32 incremental %s build, ||dP||_F = %.0e" % (("fp32", "bf16x3")[prec - 1], norm))
            xc.set_orbitals(C, occ)
            xc.build(PBE, FP64, fetch=False)
            xc.set_orbitals(C1, occ, incremental=True)
            res = xc.build(PBE, prec)
            check(res, ref, e_tol=1e-9 + k * norm, v_tol=1e-10 + k * norm, q_tol=1e-9 + k * norm,
                  what="  vs oracle:")
            assert abs(res[1] - ref[1]) <= 1e-6


# ------------------------------------------------------ 64-molecule batch
def test_batch64_bitwise_and_oracle():
    say("64 molecules: building the workloads")
    mols = [m[1:] for m in workloads.batch64()]
    say("64 molecules: %d..%d functions, %d points"
        % (min(m[0].nbas for m in mols), max(m[0].nbas for m in mols),
           sum(m[1].shape[0] for m in mols)))
    xcs = []
    for b, g, C, occ in mols:
        xc = XCBuilder(b, g, chunk_points=65536)
        xc.set_orbitals(C, occ)
        xcs.append(xc)
    try:
        say("per-molecule FP64 builds")
        singles = [xc.build(PBE, FP64) for xc in xcs]
        with Batch(xcs) as bt:
            for graph in (False, True, True):
                say("batch FP64 build, graph=%s" % graph)
                bt.build(PBE, FP64, graph=graph)
                for r, s in zip(bt.results(), singles):
                    assert np.array_equal(r[0], s[0]) and r[1] == s[1] and r[2] == s[2]
            nb = [m[0].nbas for m in mols]
            pick = sorted({int(np.argmin(nb)), int(np.argmax(nb)), int(np.argsort(nb)[32])})
            for i in pick:
                b, g, C, occ = mols[i]
                say("molecule %d: N=%d, %d points, oracle on %d threads" % (i, b.nbas, g.shape[0], NT))
                ref = oracle.xc_build(b, g, oracle.MASK_PBE, C=C, occ=occ, nthreads=NT)
                check(singles[i], ref, what="  molecule %d FP64" % i)
            # split-BF16 incremental batch builds on a new density
            Cs = [systems.perturbed_orbitals(m[2], 1e-4, seed=4000 + i) for i, m in enumerate(mols)]
            for xc, Ck, m in zip(xcs, Cs, mols):
                xc.set_orbitals(Ck, m[3], incremental=True)
            say("batch split-BF16 incremental build")
            bt.build(PBE, BF16X3, graph=True)
            mixed = bt.results()
        worst = 0.0
        for xc, Ck, m, mx in zip(xcs, Cs, mols, mixed):
            xc.set_orbitals(Ck, m[3])
            v, e, _ = xc.build(PBE, FP64)
            worst = max(worst, abs(mx[1] - e))
        say("max |dE| split-BF16 batch vs FP64: %.3e" % worst)
        assert worst <= 1e-6
    finally:
        for xc in xcs:
            xc.close()


# ----------------------------------------------- (H2O)128 + 40-atom solute
def test_water_solute_schedules_and_oracle_slices():
    say("(H2O)128 + solute: building the workload")
    els, b, g, C, occ = workloads.water_solute()
    say("(H2O)128 + solute: N=%d, %d points" % (b.nbas, g.shape[0]))
    model = default_model()
    assert model is not None
    modes = [("static", _lib.SCHED_STATIC, None), ("analytic", _lib.SCHED_ANALYTIC, None),
             ("model", _lib.SCHED_MODEL, model)]
    with XCBuilder(b, g) as xc:
        xc.set_orbitals(C, occ)
        for prec in (FP64, BF16X3):
            ref, ms = None, {}
            for name, mode, m in modes:
                xc.set_schedule(mode, m)
                out = xc.build(PBE, prec)
                ts = []
                for _ in range(2):  # timed builds after the one that plans
                    xc.build(PBE, prec, fetch=False)
                    ts.append(xc.timing()["total"])
                ms[name] = min(ts)
                say("  %s schedule, precision %d: %.1f ms" % (name, prec, ms[name]))
                if ref is None:
                    ref = out
                else:
                    assert np.array_equal(out[0], ref[0]) and out[1] == ref[1] and out[2] == ref[2]
            # what the trained network gains over counted flops and over the
            # static split (the reference's method); no schedule falls far behind
            say("  model gain: %.3fx over analytic, %.3fx over static"
                % (ms["analytic"] / ms["model"], ms["static"] / ms["model"]))
            assert max(ms.values()) <= 1.25 * min(ms.values())
    model.close()
    npts = g.shape[0]
    for lo in (npts // 7, npts // 2, (5 * npts) // 6):
        gs = np.asfortranarray(g[lo:lo + 8192])
        say("slice of 8192 points from %d: GPU FP64, oracle on %d threads" % (lo, NT))
        with XCBuilder(b, gs) as xc:
            xc.set_orbitals(C, occ)
            res = xc.build(PBE, FP64)
        ref = oracle.xc_build(b, gs, oracle.MASK_PBE, C=C, occ=occ, nthreads=NT)
        assert abs(res[2]) > 1e-3  # the slice holds density
        check(res, ref, what="  slice %d FP64" % lo)
