#!/usr/bin/env python3
"""XC-build throughput on one MI355X (grid points per second).

Workload: the (H2O)32 cluster (96 atoms, 1056 basis functions, 1.96 M
Becke-partitioned grid points, 75 radial x 302 Lebedev per atom, 160
occupied orbitals, PBE).  A step is one full XC build over the whole grid,
the basis values and gradients evaluated in it (as the reference evaluates
them in every build).
  value        : mixed-precision build run as an incremental difference-density
                 build (each step a new density, ||dP||_F = 1e-4 from the
                 last), inputs resident in HBM; device time of the build.  The
                 fastest of the FP32-MFMA and split-BF16 forms whose E_xc
                 stays within 1e-6 Eh of the FP64 build is reported.
  fp64_*       : the plain FP64 GPU build of the same workload.
  dE_vs_fp64   : E_xc after the chain of mixed incremental builds minus E_xc
                 of a full FP64 build of the final density; dV_vs_fp64 the
                 same for max |V_xc|.
  cpu_baseline : the whole FP64 build of the CPU oracle on the host's cores;
                 its result also checks the GPU FP64 build at this size
                 (gpu_fp64_vs_cpu).
  warm_cache   : the same builds with the basis values kept in HBM across
                 builds of the geometry.
  roofline     : the density contraction (the larger contraction) timed on
                 one lane: its algorithmic bytes per launch (P sub-block in
                 FP64, values and gradients in FP32) over the launch time,
                 against 8 TB/s when its intensity (flops per algorithmic
                 byte) is below the ridge of its MFMA type, else its flops
                 against the MFMA peak; the other rate is reported beside it.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "paragauss-gpl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}  # TFLOP/s, MI355X dense


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    from pgxc import workloads
    if name not in ("h2o32", "h2o"):
        raise SystemExit("unknown workload " + name)
    t = time.time()
    els, b, g, C, occ = getattr(workloads, name)(partition="gpu")
    log("grid: %d points for %d atoms in %.1f s" % (g.shape[0], len(els), time.time() - t))
    return els, b, g, C, occ


def density_chain(C, n, norm=1e-4, seed=14):
    """n orbital sets, each a perturbation of the previous with ||dP||_F = norm."""
    from pgxc import systems
    out, cur = [], C
    for k in range(n):
        cur = systems.perturbed_orbitals(cur, norm, seed=seed + k)
        out.append(cur)
    return out


def precision_errors(xc, C, occ, precs, norms=(1e-2, 1e-4, 1e-6)):
    """|dE_xc| and max|dV| of one incremental mixed build per ||dP||_F, each
    against the FP64 full build of the same density."""
    import numpy as np
    from pgxc.xc import FP64, PBE, PRECISIONS
    out = {}
    for k, norm in enumerate(norms):
        from pgxc import systems
        C1 = systems.perturbed_orbitals(C, norm, seed=100 + k)
        xc.set_orbitals(C1, occ)
        v_ref, e_ref, _ = xc.build(PBE, FP64)
        for prec in precs:
            xc.set_orbitals(C, occ)
            xc.build(PBE, FP64, fetch=False)
            xc.set_orbitals(C1, occ, incremental=True)
            v, e, _ = xc.build(PBE, PRECISIONS[prec])
            out.setdefault(prec, {})["%.0e" % norm] = {
                "dE": e - e_ref, "dV": float(np.max(np.abs(v - v_ref)))}
    return out


def schedule_ab(xc, chain, occ, prec, model, steps):
    """The mixed build under the static split, counted-flop queues and
    cost-model queues, interleaved build by build on one device; mean device
    ms per build and the model's gains."""
    import numpy as np
    from pgxc import _lib
    from pgxc.xc import PBE
    modes = [("static", _lib.SCHED_STATIC, None), ("analytic", _lib.SCHED_ANALYTIC, None),
             ("model", _lib.SCHED_MODEL, model)]
    ms = {n: [] for n, _, _ in modes}
    for r in range(max(2, steps)):
        for name, mode, m in modes:
            xc.set_schedule(mode, m)
            for k in range(2):  # the first build after a change of schedule plans
                xc.set_orbitals(chain[(2 * r + k) % len(chain)], occ, incremental=True)
                xc.build(PBE, prec, fetch=False)
            ms[name].append(xc.timing()["total"])
    out = {n: float(np.mean(v)) for n, v in ms.items()}
    out["model_gain_vs_analytic"] = out["analytic"] / out["model"]
    out["model_gain_vs_static"] = out["static"] / out["model"]
    return out


def copy_calibration(mib):
    """A plain streaming kernel of known size (y = -x over mib MiB of FP32) for
    scaling the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    import torch
    x = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    del x, y


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(b, g, C, occ, gpu_fp64=None):
    """The FP64 oracle's full build of the whole workload on the host's
    cores (occupied-orbital density, the reference's form), timed; with the
    GPU's FP64 build of the same density, the differences between the two."""
    import numpy as np
    import oracle
    nt = min(16, os.cpu_count() or 1)
    t = time.time()
    ref = oracle.xc_build(b, g, oracle.MASK_PBE, C=C, occ=occ, veclen=128, nthreads=nt)
    dt = time.time() - t
    out = {"value": g.shape[0] / dt, "unit": "grid pts/s", "cores": nt, "kind": "port",
           "cpu": cpu_model(), "seconds": dt,
           "sample": "the whole workload: one full FP64 oracle build (%d points in 128-point "
                     "batches, occupied-orbital density as the reference), %d threads"
                     % (g.shape[0], nt)}
    if gpu_fp64 is not None:
        v, e, q = gpu_fp64
        out["gpu_fp64_vs_cpu"] = {"dE": e - ref[1], "max_dV": float(np.max(np.abs(v - ref[0]))),
                                  "dQ": q - ref[2]}
    return out


def profiled_traffic(kernel):
    """HBM bytes per dispatch of a kernel, from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size),
    or None when the summary does not cover it."""
    path = os.path.join(ROOT, "profiles", "xc_build_h2o32_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_dispatch" and f[1] == kernel:
            return float(f[2])
    return None


def learned_training_set(log, max_items=1 << 18):
    """Per-item timings logged by the library on the H2O, (H2O)32 FP64 and
    (H2O)32 mixed-precision builds: features and log(us)."""
    import numpy as np
    from pgxc import XCBuilder, basis, grid, systems
    from pgxc.costmodel import target
    from pgxc.xc import BF16X3, FP32, FP64, PBE
    Xs, ys = [], []

    def run(b, g, C, occ, precs, reps, **kw):
        with XCBuilder(b, g, **kw) as xc:
            xc.set_orbitals(C, occ)
            xc.set_item_log(True)
            xc.build(PBE, FP64, fetch=False)
            for prec in precs:
                for _ in range(reps):
                    if prec != FP64:
                        xc.set_orbitals(systems.perturbed_orbitals(C, 1e-4, seed=5), occ,
                                        incremental=True)
                    xc.build(PBE, prec, fetch=False)
                    F, us = xc.item_log()
                    Xs.append(F)
                    ys.append(target(us))
    els, xyz = systems.h2o()
    b = basis.seeded_basis(els, xyz, seed=1, shell_set="small")
    g = grid.molecular_grid(els, xyz, nrad=75, nang=302)
    C, occ = systems.random_orbitals(b.nbas, 5, seed=2)
    run(b, g, C, occ, [FP64], 2)
    els, b, g, C, occ = workload("h2o32")
    for tile in (64, 96, 128):
        run(b, g, C, occ, [FP64, FP32, BF16X3], 1, tile=tile)
    X, y = np.vstack(Xs), np.concatenate(ys)
    if len(y) > max_items:  # a seeded subsample keeps training to seconds
        keep = np.sort(np.random.default_rng(17).choice(len(y), max_items, replace=False))
        X, y = X[keep], y[keep]
    log("cost-model training set: %d items" % len(y))
    return X, y


def bench_learned(args):
    """(H2O)128 + 40-atom solute: static work split vs queues balanced by
    counted flops vs queues balanced by the trained cost model."""
    import numpy as np
    from pgxc import XCBuilder, _lib, basis, grid, systems
    from pgxc.costmodel import CostModel
    from pgxc.xc import BF16X3, FP64, PBE
    X, y = learned_training_set(log)
    model = CostModel(seed=3)
    t = time.time()
    loss = model.train(X, y, steps=args.train_steps, lr=3e-3)
    log("cost model trained on the GPU: %d steps, loss %.4f (standardised), %.1f s"
        % (args.train_steps, loss, time.time() - t))
    if args.save_model:
        P, nrm = model.params()
        np.savez(args.save_model, params=P, norm=nrm, items=len(y), steps=args.train_steps,
                 loss=loss)
        log("cost model saved to", args.save_model)
    from pgxc import workloads
    t = time.time()
    els, b, g, C, occ = workloads.water_solute()
    log("grid: %d points for %d atoms in %.1f s" % (g.shape[0], len(els), time.time() - t))
    npts, n = g.shape[0], b.nbas
    xc = XCBuilder(b, g)
    xc.set_orbitals(C, occ)
    log("layout:", xc.layout())
    modes = [("static", _lib.SCHED_STATIC, None), ("analytic", _lib.SCHED_ANALYTIC, None),
             ("model", _lib.SCHED_MODEL, model)]
    res, ref = {}, {}
    for prec_name, prec in (("fp64", FP64), ("bf16x3", BF16X3)):
        for name, mode, m in modes:
            xc.set_schedule(mode, m)
            t = time.time()
            xc.build(PBE, prec, fetch=False)  # plan (features, predictions) + warm-up
            tplan = time.time() - t
            ms = []
            for _ in range(args.steps):
                xc.build(PBE, prec, fetch=False)
                ms.append(xc.timing()["total"])
            out = xc.result()
            if prec_name not in ref:
                ref[prec_name] = out
            same = bool(np.array_equal(out[0], ref[prec_name][0]) and out[1] == ref[prec_name][1])
            res["%s/%s" % (prec_name, name)] = {"ms": min(ms), "pts_per_s": npts / (min(ms) * 1e-3),
                                               "bit_identical_to_static": same,
                                               "first_build_s": tplan}
            log("%s %s: %.1f ms (first build %.1f s) identical=%s"
                % (prec_name, name, min(ms), tplan, same))
    st = xc.stats()
    xc.close()
    model.close()
    best = res["bf16x3/model"]
    out = {
        "metric": "XC-build grid pts/s, learned work queues vs static split, (H2O)128 + 40-atom solute",
        "value": best["pts_per_s"], "unit": "grid pts/s", "n_gpus": 1, "steps": args.steps,
        "higher_is_better": True, "dtype": "f64+bf16x3", "data": "synthetic",
        "config": {"workload": "(H2O)128 + 40-atom C/H/N/O solute, %d basis functions, %d grid "
                               "points, PBE, full builds" % (n, npts)},
        "schedules": res,
        "model_gain_vs_static": {p: res[p + "/static"]["ms"] / res[p + "/model"]["ms"]
                                 for p in ("fp64", "bf16x3")},
        "model_gain_vs_analytic": {p: res[p + "/analytic"]["ms"] / res[p + "/model"]["ms"]
                                   for p in ("fp64", "bf16x3")},
        "fp64_static_pts_per_s": res["fp64/static"]["pts_per_s"],
        "screened_flops_per_build": st["flops"], "dense_flops_per_build": 4.0 * npts * n * n + 16.0 * npts * n,
        "training_items": int(len(y)),
    }
    print(json.dumps(out), flush=True)


def bench_batch64(args):
    """64 seeded C/H/N/O molecules: per-molecule FP64 builds vs one batched,
    hipGraph-replayed build (FP64, and split-BF16 incremental)."""
    import numpy as np
    from pgxc import Batch, XCBuilder, autotune, basis, grid, systems
    from pgxc.xc import BF16X3, FP64, PBE
    from pgxc import workloads
    t = time.time()
    mols = [m[1:] for m in workloads.batch64()]
    npts = sum(m[1].shape[0] for m in mols)
    log("64 molecules, %d..%d basis functions, %d grid points, %.1f s"
        % (min(m[0].nbas for m in mols), max(m[0].nbas for m in mols), npts, time.time() - t))
    t = time.time()
    xcs = []
    for b, g, C, occ in mols:
        tu = autotune.tune(b, g, C, occ, sample_points=8192, reps=1, mixed=BF16X3)
        # every context of a batch holds its own work chunk while the batch
        # runs, so chunks of 64 Ki points keep 64 of them within HBM
        xc = XCBuilder(b, g, tile=tu["tile"], variants=(tu["variant_fp64"], tu["variant_fp32"]),
                       chunk_points=65536)
        xc.set_orbitals(C, occ)
        xcs.append(xc)
    log("autotuned and created 64 contexts in %.1f s" % (time.time() - t))
    # per-molecule FP64 launches
    for xc in xcs:
        xc.build(PBE, FP64, fetch=False)
    per = []
    for _ in range(args.steps):
        t0 = time.time()
        for xc in xcs:
            xc.build(PBE, FP64, fetch=False)
        per.append(time.time() - t0)
    ref = [xc.result() for xc in xcs]
    with Batch(xcs) as bt:
        bt.build(PBE, FP64, graph=True)
        b64 = []
        for _ in range(args.steps):
            t0 = time.time()
            bt.build(PBE, FP64, graph=True)
            b64.append(time.time() - t0)
        same = all(np.array_equal(x.result()[0], r[0]) for x, r in zip(xcs, ref))
        # mixed precision: incremental split-BF16 builds from the FP64 state
        # along a chain of densities, ||dP||_F = 1e-4 per step and molecule
        cur = [m[2] for m in mols]
        bx3 = []
        for step in range(args.steps + 1):  # the first (untimed) captures the graph
            cur = [systems.perturbed_orbitals(Ck, 1e-4, seed=4000 + 97 * step + i)
                   for i, Ck in enumerate(cur)]
            for xc, Ck, m in zip(xcs, cur, mols):
                xc.set_orbitals(Ck, m[3], incremental=True)
            t0 = time.time()
            bt.build(PBE, BF16X3, graph=True)
            if step:
                bx3.append(time.time() - t0)
        mixed = [x.result() for x in xcs]
        chains = cur
    dE = []
    for xc, Ck, m, mx in zip(xcs, chains, mols, mixed):
        xc.set_orbitals(Ck, m[3])
        v, e, _ = xc.build(PBE, FP64)
        dE.append(abs(mx[1] - e))
    for xc in xcs:
        xc.close()
    t_per, t_b64, t_bx3 = min(per), min(b64), min(bx3)
    out = {
        "metric": "XC-build grid pts/s, 64-molecule batch in one hipGraph vs per-molecule FP64 launches",
        "value": npts / t_bx3, "unit": "grid pts/s", "n_gpus": 1, "steps": args.steps,
        "higher_is_better": True, "dtype": "f64+bf16x3", "data": "synthetic",
        "config": {"workload": "64 seeded C/H/N/O molecules (3-60 atoms), %d grid points, "
                               "50-75 x 194/302 grids, PBE" % npts},
        "per_molecule_fp64_s": t_per, "batch_fp64_graph_s": t_b64, "batch_bf16x3_graph_s": t_bx3,
        "batch_fp64_bit_identical": same,
        "speedup_batch_fp64": t_per / t_b64, "speedup_batch_bf16x3": t_per / t_bx3,
        "max_abs_dE_bf16x3_vs_fp64": max(dE),
    }
    print(json.dumps(out), flush=True)


def bench_open_shell(args):
    """(H2O)32 with an open shell (ispin = 2): alpha 160, beta 159 occupied
    orbitals.  FP64 full builds and split-BF16 incremental builds along a
    chain of densities, against the closed-shell build of the same grid."""
    import numpy as np
    from pgxc import XCBuilder, autotune, systems
    from pgxc.xc import BF16X3, FP64, PBE
    els, b, g, C, occ = workload("h2o32")
    npts = g.shape[0]
    Cb = systems.perturbed_orbitals(C, 1e-2, seed=21)
    Cs = np.stack([C, Cb], axis=2)
    occs = np.stack([occ, np.concatenate([occ[:-1], [0.0]])], axis=1)
    tu = None if args.no_tune else autotune.tune(b, g, C, occ, mixed=BF16X3, log=log)
    xc = XCBuilder(b, g, tile=tu["tile"] if tu else 0,
                   variants=(tu["variant_fp64"], tu["variant_fp32"]) if tu else (-1, -1))
    headline_schedule(xc)

    def timed(prec, incremental_chain=None):
        out = []
        for i in range(args.warmup + args.steps):
            if incremental_chain is not None:
                xc.set_orbitals(incremental_chain[i], occs, incremental=True)
            xc.build(PBE, prec, fetch=False)
            if i >= args.warmup:
                out.append(xc.timing()["total"])
        return sum(out) / len(out)

    xc.set_orbitals(C, occ)
    t_closed = timed(FP64)
    xc.set_orbitals(Cs, occs)
    t64 = timed(FP64)
    chain, cur = [], Cs
    for k in range(args.warmup + args.steps):
        cur = np.stack([systems.perturbed_orbitals(cur[:, :, s], 1e-4, seed=40 + 2 * k + s)
                        for s in range(2)], axis=2)
        chain.append(cur)
    xc.set_orbitals(Cs, occs)
    xc.build(PBE, FP64, fetch=False)
    tx3 = timed(BF16X3, chain)
    v, e, _ = xc.result()
    xc.set_orbitals(chain[-1], occs)
    v64, e64, _ = xc.build(PBE, FP64)
    xc.close()
    log("open shell: closed-shell fp64 %.2f ms, open-shell fp64 %.2f ms, bf16x3 %.2f ms"
        % (t_closed, t64, tx3))
    out = {
        "metric": "XC-build grid pts/s, open shell (ispin = 2), mixed vs FP64 GPU",
        "value": npts / (tx3 * 1e-3), "unit": "grid pts/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": tx3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64+bf16x3", "data": "synthetic",
        "config": {"workload": "(H2O)32 open shell, alpha 160 / beta 159 occupied, %d grid points, "
                               "N = %d, PBE" % (npts, b.nbas)},
        "closed_shell_fp64_ms": t_closed, "open_shell_fp64_ms": t64, "open_shell_bf16x3_ms": tx3,
        "speedup_vs_fp64": t64 / tx3, "abs_dE_bf16x3_vs_fp64": abs(e - e64),
        "max_abs_dV_bf16x3_vs_fp64": float(np.max(np.abs(v - v64))),
    }
    print(json.dumps(out), flush=True)


def bench_post_scf(args):
    """(H2O)32 after an FP64 build: post-SCF energies per term (FP64 and
    split-BF16), the basis-centre XC gradient and the full gradient with the
    grid moving with the atoms."""
    import numpy as np
    from pgxc import XCBuilder, grid
    from pgxc.xc import BF16X3, FP64, PBE
    els, b, g, C, occ = workload("h2o32")
    sizes = [len(grid.atomic_grid(c, 75, 302, grid.RADIUS[e])[0]) for e, c in zip(els, b.atom_xyz)]
    owner = np.repeat(np.arange(len(els)), sizes)
    xc = XCBuilder(b, g, tile=96)
    xc.set_orbitals(C, occ)
    xc.build(PBE, FP64, fetch=False)
    out = {}
    for prec, name in ((FP64, "fp64"), (BF16X3, "bf16x3")):
        t = []
        for _ in range(args.warmup + args.steps):
            t0 = time.time()
            xc.energy(PBE, prec)
            t.append(time.time() - t0)
        out["energy_%s_ms" % name] = 1e3 * min(t[args.warmup:])
    xc.build(PBE, FP64, fetch=False)
    t0 = time.time()
    g1 = xc.gradient()
    out["gradient_s"] = time.time() - t0
    t0 = time.time()
    g2 = xc.gradient_grid(owner)
    out["gradient_grid_s"] = time.time() - t0
    xc.close()
    out["gradient_grid_translation_residual"] = float(np.abs(g2.sum(0)).max())
    out["gradient_max"] = float(np.abs(g2).max())
    log("post-SCF:", out)
    print(json.dumps({"metric": "post-SCF XC energies and gradients, (H2O)32",
                      "value": out["gradient_grid_s"], "unit": "s per full gradient",
                      "n_gpus": 1, "higher_is_better": False, "dtype": "f64", "data": "synthetic",
                      "config": {"workload": "(H2O)32, N = %d, %d grid points, PBE" % (b.nbas, g.shape[0])},
                      **out}), flush=True)


def headline_schedule(xc):
    """The headline's contraction schedule: queues ordered by the shipped cost
    model, or by counted flops where no model ships."""
    from pgxc import _lib
    from pgxc.costmodel import default_model
    model = default_model()
    xc.set_schedule(_lib.SCHED_MODEL if model is not None else _lib.SCHED_ANALYTIC, model)


def bench_profile(args):
    """The kernels the headline's roofline is taken from, alone, for rocprofv3
    (kernel trace and traffic counters): a copy kernel of known size (counter
    calibration), then (H2O)32 builds on one lane, with the basis evaluated and
    with it cached, FP64 full builds and split-BF16 incremental builds."""
    from pgxc import XCBuilder, systems
    from pgxc.xc import BF16X3, FP64, PBE
    copy_calibration(1024)
    els, b, g, C, occ = workload("h2o32")
    for cache in (False, True):
        xc = XCBuilder(b, g, tile=96, variants=(1, 1), basis_cache=cache)
        xc.set_lanes(1)
        headline_schedule(xc)
        xc.set_orbitals(C, occ)
        for prec in (FP64, BF16X3):
            for i in range(args.steps):
                xc.set_orbitals(systems.perturbed_orbitals(C, 1e-4, seed=50 + i), occ,
                                incremental=prec != FP64)
                xc.build(PBE, prec, fetch=False)
                log("basis %s, %s, build %d:" % ("cached" if cache else "evaluated",
                                                  "fp64" if prec == FP64 else "bf16x3", i),
                    {k: round(v, 3) for k, v in xc.timing().items()})
        xc.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="h2o32",
                    help="h2o32 (headline), h2o, batch64 (64-molecule batch), "
                         "learned ((H2O)128 + solute, cost-model schedules), "
                         "h2o32-open ((H2O)32 open shell), "
                         "post-scf ((H2O)32 energies per term and XC gradients), "
                         "h2o32-profile (one-lane builds for rocprofv3)")
    ap.add_argument("--precision", default="auto", choices=["auto", "fp32", "bf16x3"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the tile autotuner")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a streaming kernel over MIB MiB (counter calibration)")
    ap.add_argument("--train-steps", type=int, default=2000, help="cost-model training steps")
    ap.add_argument("--save-model", default="", help="learned: write the trained model here (.npz)")
    ap.add_argument("--schedule", default="auto", choices=["auto", "static", "analytic", "model"],
                    help="contraction work schedule: static split, queues balanced by counted "
                         "flops, or by the shipped cost model's predictions (auto: the model "
                         "when shipped, else analytic)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(the multi-GPU sharding is a design note only)" % args.gpus)
        sys.exit(2)
    if args.workload == "learned":
        return bench_learned(args)
    if args.workload == "batch64":
        return bench_batch64(args)
    if args.workload == "h2o32-open":
        return bench_open_shell(args)
    if args.workload == "post-scf":
        return bench_post_scf(args)
    if args.workload == "h2o32-profile":
        return bench_profile(args)

    import numpy as np
    from pgxc import XCBuilder
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    from pgxc.xc import FP64, PBE, PRECISIONS

    els, b, g, C, occ = workload(args.workload)
    npts, n = g.shape[0], b.nbas
    precs = ["fp32", "bf16x3"] if args.precision == "auto" else [args.precision]
    kw = {}
    if not args.no_tune:
        from pgxc import autotune
        t0 = time.time()
        tuned = autotune.tune(b, g, C, occ, mixed=PRECISIONS["fp32" if precs == ["fp32"] else "bf16x3"],
                              log=log)
        log("autotune: %s in %.1f s" % ({k: v for k, v in tuned.items() if k != "table"},
                                         time.time() - t0))
        kw = dict(tile=tuned["tile"], variants=(tuned["variant_fp64"], tuned["variant_fp32"]))
    from pgxc import _lib
    from pgxc.costmodel import default_model
    sched, model = args.schedule, default_model()
    if sched == "auto":
        sched = "model" if model is not None else "analytic"
    if sched == "model" and model is None:
        raise SystemExit("bench.py: no shipped cost model (pgxc/cost_model_mi355x.npz)")
    sched_mode = {"static": _lib.SCHED_STATIC, "analytic": _lib.SCHED_ANALYTIC,
                  "model": _lib.SCHED_MODEL}[sched]
    log("schedule:", sched)
    chain = density_chain(C, args.warmup + args.steps)

    def timed(xc, label):
        """FP64 full builds of C, then mixed incremental builds along the chain
        of densities, each chain's end checked against an FP64 full build."""
        xc.set_orbitals(C, occ)
        for _ in range(args.warmup):
            xc.build(PBE, FP64, fetch=False)
        t64, st64 = [], []
        for _ in range(args.steps):
            xc.build(PBE, FP64, fetch=False)
            tm = xc.timing()
            t64.append(tm["total"])
            st64.append(tm)
        res64 = xc.result()
        log("%s fp64: %.2f ms/build  stages %s" % (label, sum(t64) / len(t64), st64[-1]))
        runs = {}
        for prec in precs:
            xc.set_orbitals(C, occ)
            xc.build(PBE, FP64, fetch=False)  # FP64 state the chain starts from
            times, stages, walls = [], [], []
            for i, Ck in enumerate(chain):
                t0 = time.time()
                xc.set_orbitals(Ck, occ, incremental=True)  # host orbitals -> device, P formed there
                xc.build(PBE, PRECISIONS[prec], fetch=True)  # packed V, E_xc, charge back to the host
                walls.append(time.time() - t0)
                if i >= args.warmup:
                    tm = xc.timing()
                    times.append(tm["total"])
                    stages.append(tm)
            v_mx, e_mx, _ = xc.result()
            xc.set_orbitals(chain[-1], occ)
            v_ref, e_ref, _ = xc.build(PBE, FP64)
            runs[prec] = dict(ms=sum(times) / len(times), st=stages[-1], dE=e_mx - e_ref,
                              wall_ms=1e3 * min(walls[args.warmup:] or walls),
                              dV=float(np.max(np.abs(v_mx - v_ref))), E=e_ref)
            log("%s %s: %.2f ms/build dE=%.3e dV=%.3e stages %s" % (
                label, prec, runs[prec]["ms"], runs[prec]["dE"], runs[prec]["dV"], stages[-1]))
        ok = [p for p in precs if abs(runs[p]["dE"]) <= 1e-6] or precs
        best = min(ok, key=lambda p: runs[p]["ms"])
        return dict(ms64=sum(t64) / len(t64), st64=st64[-1], res64=res64, runs=runs, best=best)

    def one_lane(xc, prec):
        """Per-stage times of builds on one lane, where each launch has the
        device to itself (the timed builds overlap two chunks' launches)."""
        xc.set_lanes(1)
        one = []
        for _ in range(max(1, args.steps)):
            xc.set_orbitals(chain[-1], occ, incremental=True)
            xc.build(PBE, PRECISIONS[prec], fetch=False)
            one.append(xc.timing())
        xc.set_lanes(2)
        return {k: float(np.mean([t[k] for t in one])) for k in one[0]}, len(one)

    # ---- builds that evaluate the basis every time, as the reference does
    # (orbital_calculate per batch, modules/xc_hamiltonian.f90:1365)
    xc = XCBuilder(b, g, basis_cache=False, **kw)
    xc.set_schedule(sched_mode, model)
    layout = xc.layout()
    log("layout:", layout)
    cold = timed(xc, "basis evaluated")
    best = cold["best"]
    st1, nlaunch = one_lane(xc, best)
    log("%s on one lane (basis evaluated): %.2f ms/build stages %s" % (best, st1["total"], st1))
    stats = xc.stats()
    sched_ab = schedule_ab(xc, chain, occ, PRECISIONS[best], model, args.steps) if model else None
    if sched_ab:
        log("schedules (interleaved, ms per build):", sched_ab)
    xc.set_schedule(sched_mode, model)
    xc.close()

    # ---- SCF-cycle builds with the basis values kept in HBM across builds
    xc = XCBuilder(b, g, **kw)
    xc.set_schedule(sched_mode, model)
    warm = timed(xc, "basis cached")
    st1w, _ = one_lane(xc, warm["best"])
    # work the contraction kernels execute (gathered runs of 8 functions,
    # whole function tiles): counted from the item log of one more build of
    # each mixed form
    executed = {}
    xc.set_item_log(True)
    for prec in precs:
        xc.set_orbitals(chain[-1], occ, incremental=True)
        xc.build(PBE, PRECISIONS[prec], fetch=False)
        F, _ = xc.item_log()
        k = np.exp2(16.0 * F[:, 5]) - 1.0
        # density item: 128 points x a pair of function tiles (both computed
        # by the split-BF16 kernel); potential item: function tile x tile
        tile = F[:, 4] * 128.0
        rows = np.where(F[:, 0] == 0, 128.0, tile)
        cols = np.where(F[:, 0] == 0, 2.0 * tile, tile)
        fl = 2.0 * rows * cols * k
        executed[prec] = {"density": float(fl[F[:, 0] == 0].sum()),
                          "potential": float(fl[F[:, 0] == 1].sum())}
    xc.set_item_log(False)
    perr = precision_errors(xc, C, occ, precs)
    log("precision errors per ||dP||_F:", perr)
    xc.close()

    R, Rw = cold["runs"][best], warm["runs"][warm["best"]]
    ms64, msmx = cold["ms64"], R["ms"]
    # dominant kernel of the mixed build and its roofline.  Algorithmic flops
    # of one contraction: half of pgxc_last_stats' sum over 128-point tiles of
    # 4 B n_b^2 + 16 B n_b (n_b: functions not cut off on the tile; the 16 B n_b
    # elementwise share is ~1 %); the dense equivalent 2 npts n^2 beside it.
    # The basis stage (k_basis) is shorter per chunk than the density
    # contraction (kernel trace of --workload h2o32-profile), so the dominant
    # kernel is the larger contraction; the basis stage gets its own HBM
    # roofline below.
    st = st1
    dom = max(("density", "potential"), key=lambda k: st[k])
    tf = layout["tile"]
    peak = PEAK["fp32"] if best == "fp32" else PEAK["bf16"] / 3.0
    flops = stats["flops"] / 2.0
    achieved = flops / (st[dom] * 1e-3) / 1e12
    if best == "bf16x3":
        kname = "k_%s_bf16x3<%d%s>" % (dom, tf, ",true" if dom == "density" else "")
    else:
        v = layout["variant_fp32"]
        kname = "k_%s<float,%d,%d%s>" % (dom, tf, v, ",true" if dom == "density" else "")
    mfma = {"achieved": achieved, "peak": peak,
            "unit": "TFLOP/s" + (" (FP32-equivalent: BF16 peak / 3 passes)" if best == "bf16x3" else ""),
            "frac": achieved / peak,
            "achieved_dense_equivalent": 2.0 * npts * n * n / (st[dom] * 1e-3) / 1e12,
            "achieved_executed": executed[best][dom] / (st[dom] * 1e-3) / 1e12,
            "flops_per_build_executed": executed[best],
            "flops_per_build_algorithmic": stats["flops"],
            "flops_per_build_dense": 4.0 * npts * n * n + 16.0 * npts * n}
    # Algorithmic bytes of the contraction (per batch b of B_b points and n_b
    # live functions): the density reads the P sub-block (8 n_b^2, FP64) and
    # the stored values and gradients (4 planes of 4-byte elements,
    # 16 B_b n_b); the potential writes the V partial (8 n_b^2) and reads the
    # stored values (4 B_b n_b).  Against these bytes the contraction's
    # arithmetic intensity sits below the ridge of its MFMA type (peak
    # flops / 8 TB/s), which makes HBM its roofline bound.
    abytes = (8.0 * stats["n2"] + (16.0 if dom == "density" else 4.0) * stats["bn"])
    intensity = flops / abytes
    ridge = peak / 8.0
    if intensity < ridge:
        ach_bw = abytes / (st[dom] * 1e-3) / 1e12
        roof = {"bound": "hbm", "kernel": kname, "achieved": ach_bw, "peak": 8.0, "unit": "TB/s",
                "frac": ach_bw / 8.0, "traffic": profiled_traffic(kname)}
    else:
        roof = {"bound": "mfma", "kernel": kname, "achieved": achieved, "peak": peak,
                "unit": mfma["unit"], "frac": achieved / peak, "traffic": profiled_traffic(kname)}
    roof.update({"algorithmic_bytes_per_launch": abytes / layout["chunks"],
                 "algorithmic_flops_per_launch": flops / layout["chunks"],
                 "intensity_flop_per_byte": intensity, "ridge_flop_per_byte": ridge, "mfma": mfma})
    # basis stage (k_basis) against the HBM roof: FP32 values and
    # gradients of the live functions written once, and the grid read
    bytes_basis = 32.0 * npts + 16.0 * stats["bn"]
    ach_b = bytes_basis / (st["basis"] * 1e-3) / 1e12
    # the basis kernel is built for the largest angular momentum of the basis
    kbasis = "k_basis<true,float,%d>" % max(1, max(sh.l for sh in b.shells))
    roof_basis = {"bound": "hbm", "kernels": kbasis,
                  "stage_ms_one_lane": st["basis"], "bytes_per_build_algorithmic": bytes_basis,
                  "achieved": ach_b, "peak": 8.0, "unit": "TB/s", "frac": ach_b / 8.0,
                  "traffic": profiled_traffic(kbasis)}
    roof["timed"] = "one-lane builds with the basis evaluated (%d launches)" % (nlaunch * layout["chunks"])
    roof["traffic_unit"] = ("bytes per dispatch (one grid chunk), rocprofv3 FETCH_SIZE + WRITE_SIZE "
                            "scaled by a copy kernel")
    roof["dispatch_ms"] = st[dom] / layout["chunks"]
    if roof["traffic"]:  # measured bytes (rocprofv3 counters) at the live dispatch time
        hbm = roof["traffic"] / (roof["dispatch_ms"] * 1e-3) / 1e12
        roof["traffic_rate"] = {"achieved": hbm, "peak": 8.0, "unit": "TB/s", "frac": hbm / 8.0,
                                "traffic_over_algorithmic": roof["traffic"] / roof["algorithmic_bytes_per_launch"]}
    # the warm-cache build's dominant contraction, for comparison with the
    # kernel work of the previous measurements
    domw = "density" if st1w["density"] >= st1w["potential"] else "potential"
    roof_warm = {"kernel": domw, "stage_ms_one_lane": st1w[domw],
                 "achieved": stats["flops"] / 2.0 / (st1w[domw] * 1e-3) / 1e12,
                 "peak": PEAK["fp32"] if warm["best"] == "fp32" else PEAK["bf16"] / 3.0}
    roof_warm["frac"] = roof_warm["achieved"] / roof_warm["peak"]
    wl = {"h2o32": "(H2O)32", "h2o": "H2O"}[args.workload]
    out = {
        "metric": "XC-build grid pts/s, mixed vs FP64 GPU at |dE_xc|<=1e-6 Eh; %MFMA/HBM peak",
        "value": npts / (msmx * 1e-3),
        "unit": "grid pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": msmx,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+" + best,
        "data": "synthetic",
        "config": {"workload": "%s PBE XC build, %d basis functions, %d grid points, %d occupied, "
                               "basis values and gradients evaluated in every build (as the "
                               "reference), incremental dP (||dP||_F=1e-4 per step), %s contractions"
                               % (wl, n, npts, C.shape[1], best),
                   "global_batch": npts, "seq_len": 128, "parallelism": "single GPU"},
        "fp64_pts_per_s": npts / (ms64 * 1e-3),
        "fp64_ms_per_step": ms64,
        "speedup_vs_fp64": ms64 / msmx,
        "dE_vs_fp64": R["dE"],
        "dV_vs_fp64": R["dV"],
        "E_xc": R["E"],
        "wall_pts_per_s_incl_host": npts / (R["wall_ms"] * 1e-3),
        "mixed_forms": {p: {"ms_per_step": cold["runs"][p]["ms"],
                            "pts_per_s": npts / (cold["runs"][p]["ms"] * 1e-3),
                            "dE_vs_fp64": cold["runs"][p]["dE"], "dV_vs_fp64": cold["runs"][p]["dV"]}
                        for p in precs},
        "warm_cache": {"what": "SCF-cycle builds with the basis values and gradients kept in HBM "
                               "across builds (same geometry), not evaluated in the timed builds",
                       "pts_per_s": npts / (Rw["ms"] * 1e-3), "ms_per_step": Rw["ms"],
                       "precision": warm["best"], "fp64_ms_per_step": warm["ms64"],
                       "fp64_pts_per_s": npts / (warm["ms64"] * 1e-3),
                       "speedup_vs_fp64": warm["ms64"] / Rw["ms"], "dE_vs_fp64": Rw["dE"],
                       "stages_ms": {k: round(float(v), 3) for k, v in Rw["st"].items()},
                       "stages_ms_one_lane": {k: round(float(v), 3) for k, v in st1w.items()},
                       "dominant_contraction": roof_warm},
        "schedule": sched,
        "schedules_ab": sched_ab,
        "precision_errors": perr,
        "layout": {k: int(v) for k, v in layout.items()},
        "stages_ms": {k: round(float(v), 3) for k, v in R["st"].items()},
        "stages_ms_one_lane": {k: round(float(v), 3) for k, v in st1.items()},
        "fp64_stages_ms": {k: round(float(v), 3) for k, v in cold["st64"].items()},
        "algorithmic_per_build": {"flops": stats["flops"], "bytes": stats["bytes"],
                                  "exps": stats["exps"]},
        "roofline": roof,
        "roofline_basis_stage": roof_basis,
    }
    if args.workload == "h2o":
        # a build this small is latency bound, so only its HBM fraction is
        # reported: the algorithmic bytes of the whole build over its time
        ach = stats["bytes"] / (msmx * 1e-3) / 1e12
        out["roofline"] = {"bound": "hbm", "kernel": "whole build", "achieved": ach, "peak": 8.0,
                           "unit": "TB/s", "frac": ach / 8.0, "traffic": None,
                           "bytes_per_build_algorithmic": stats["bytes"]}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(b, g, C, occ, cold["res64"])
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
