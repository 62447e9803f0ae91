"""The learned cost model and the schedules it drives.

* The HIP forward / backward / training kernels against the CPU FP64 oracle
  (oracle/cost_mlp.py, itself pinned by finite differences).
* Item logging: one record per contraction work item of the last build.
* Schedules (static walk, queues balanced by counted flops, queues balanced by
  the trained model) never change a result bit, for one molecule and batches.
* What the network gains: on held-out logged items it predicts item times
  better than counted flops (the analytic cost) can.
"""
import numpy as np
import pytest

from oracle import cost_mlp
from pgxc import Batch, XCBuilder, _lib, basis, grid, systems
from pgxc.costmodel import CostModel, target
from pgxc.xc import BF16X3, FP32, FP64, PBE

pytestmark = pytest.mark.gpu
F = _lib.NFEATURES
IDENT = np.concatenate([np.zeros(F), np.ones(F), [0.0, 1.0]])


def rand_problem(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, F))
    t = rng.normal(size=n)
    P = rng.normal(scale=0.3, size=cost_mlp.nparams(F))
    return P, X, t


def test_gradient_matches_oracle():
    P, X, t = rand_problem(10000, 3)  # three partial-sum chunks of 4096
    with CostModel(F, seed=1) as m:
        m.set_params(P, IDENT)
        g, loss = m.gradient(X, t)
    gr = cost_mlp.gradient(P, X, t)
    err = np.max(np.abs(g - gr)) / np.max(np.abs(gr))
    print("max rel gradient error %.3e" % err, flush=True)
    assert err <= 1e-12
    assert abs(loss - cost_mlp.loss(P, X, t)) <= 1e-12 * cost_mlp.loss(P, X, t)


def test_predict_matches_oracle():
    P, X, _ = rand_problem(777, 4)
    with CostModel(F, seed=1) as m:
        m.set_params(P, IDENT)
        y = m.predict(X)
    yr, _, _ = cost_mlp.forward(P, X)
    assert np.max(np.abs(y - yr)) <= 1e-12 * np.max(np.abs(yr))


def test_training_converges_and_is_reproducible():
    rng = np.random.default_rng(5)
    X = rng.uniform(0, 2, size=(3000, F))
    y = np.log1p(3.0 * X[:, 5] * X[:, 4]) + 0.5 * X[:, 0]  # a smooth cost law
    out = []
    for _ in range(2):
        with CostModel(F, seed=7) as m:
            l0 = m.train(X, y, steps=0)
            l1 = m.train(X, y, steps=300, lr=1e-2)
            out.append((l0, l1, m.params()[0]))
    (l0, l1, p1), (_, l1b, p2) = out
    print("loss %.4e -> %.4e" % (l0, l1), flush=True)
    assert l1 < 0.05 * l0
    assert l1 == l1b and np.array_equal(p1, p2)


def molecule_case():
    """Two waters far apart: screening leaves very uneven work per item."""
    e1, x1 = systems.h2o()
    e2, x2 = systems.h2o(center=(30.0, 5.0, 0.0))
    els, xyz = list(e1) + list(e2), np.vstack([x1, x2])
    b = basis.seeded_basis(els, xyz, seed=31, shell_set="medium")
    g = grid.molecular_grid(els, xyz, nrad=50, nang=302)
    C, occ = systems.random_orbitals(b.nbas, 10, seed=32)
    return b, g, C, occ


@pytest.fixture(scope="module")
def case():
    return molecule_case()


def test_item_log_covers_every_item(case):
    b, g, C, occ = case
    with XCBuilder(b, g, chunk_points=8192) as xc:
        xc.set_orbitals(C, occ)
        xc.set_item_log(True)
        xc.build(PBE, FP64, fetch=False)
        Fm, us = xc.item_log()
    assert Fm.shape[0] > 0 and Fm.shape == (us.shape[0], F)
    assert np.all(us > 0)
    assert set(np.unique(Fm[:, 0])) == {0.0, 1.0}       # both contractions
    assert np.all(Fm[:, 1] == 1.0)                       # FP64 one-hot
    assert np.all(Fm[:, 5] <= Fm[:, 6] + 1e-12)          # screened K <= dense K
    assert np.any(Fm[:, 5] < Fm[:, 6])                   # screening removed work


def collect_items(b, g, C, occ, precs=(FP64, FP32, BF16X3), reps=2):
    Xs, ys = [], []
    with XCBuilder(b, g, chunk_points=8192) as xc:
        xc.set_orbitals(C, occ)
        xc.set_item_log(True)
        for prec in precs:
            for _ in range(reps):
                xc.build(PBE, prec, fetch=False)
                Fm, us = xc.item_log()
                Xs.append(Fm)
                ys.append(target(us))
    return np.vstack(Xs), np.concatenate(ys)


@pytest.fixture(scope="module")
def trained(case):
    b, g, C, occ = case
    X, y = collect_items(b, g, C, occ)
    m = CostModel(F, seed=11)
    m.train(X, y, steps=600, lr=5e-3)
    yield m, X, y
    m.close()


@pytest.mark.parametrize("prec", [FP64, BF16X3])
def test_schedules_do_not_change_bits(case, trained, prec):
    b, g, C, occ = case
    m = trained[0]
    out = []
    for mode in (_lib.SCHED_STATIC, _lib.SCHED_ANALYTIC, _lib.SCHED_MODEL):
        with XCBuilder(b, g, chunk_points=8192) as xc:
            xc.set_schedule(mode, m if mode == _lib.SCHED_MODEL else None)
            xc.set_orbitals(C, occ)
            out.append(xc.build(PBE, prec))
    for h, e, q in out[1:]:
        assert np.array_equal(h, out[0][0]) and e == out[0][1] and q == out[0][2]


def test_batch_schedules_do_not_change_bits(case, trained):
    b, g, C, occ = case
    m = trained[0]
    res = []
    for mode in (_lib.SCHED_STATIC, _lib.SCHED_MODEL):
        xcs = [XCBuilder(b, g, chunk_points=8192), XCBuilder(b, g[: g.shape[0] // 2], chunk_points=4096)]
        for xc in xcs:
            xc.set_orbitals(C, occ)
        with Batch(xcs) as bt:
            bt.set_schedule(mode, m if mode == _lib.SCHED_MODEL else None)
            for graph in (False, True, True):
                bt.build(PBE, BF16X3, graph=graph)
                res.append([x.result() for x in xcs])
        for xc in xcs:
            xc.close()
    for r in res[1:]:
        for (h, e, q), (h0, e0, q0) in zip(r, res[0]):
            assert np.array_equal(h, h0) and e == e0 and q == q0


def test_network_beats_counted_flops(case, trained):
    """Held-out items of a fresh build: the network's error in log(time)
    against the best affine law in log(counted flops), per precision."""
    b, g, C, occ = case
    m, X, y = trained
    Xt, yt = collect_items(b, g, C, occ, reps=1)
    def flops(Z):  # 2 x (tile rows) x (tile columns) x K left after screening
        k = np.exp2(16.0 * Z[:, 5]) - 1.0
        return 2.0 * np.where(Z[:, 0] == 0, 128.0, Z[:, 4] * 128) * Z[:, 4] * 128 * k
    A = np.column_stack([np.ones(len(X)), np.log1p(flops(X)), X[:, 1:4]])
    coef, *_ = np.linalg.lstsq(A, y, rcond=None)
    At = np.column_stack([np.ones(len(Xt)), np.log1p(flops(Xt)), Xt[:, 1:4]])
    e_count = np.sqrt(np.mean((At @ coef - yt) ** 2))
    e_net = np.sqrt(np.mean((m.predict(Xt) - yt) ** 2))
    print("held-out rms error of log(item time): network %.4f, counted flops %.4f"
          % (e_net, e_count), flush=True)
    assert e_net < e_count
