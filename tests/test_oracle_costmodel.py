"""The cost model's CPU oracle (oracle/cost_mlp.py), pinned without a GPU:
its back-propagated gradient against central finite differences of the loss,
and the forward pass against an explicit per-sample loop."""
import numpy as np

from oracle import cost_mlp


def problem(seed=0, n=40, F=9):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, F))
    t = rng.normal(size=n)
    P = rng.normal(scale=0.3, size=cost_mlp.nparams(F))
    return P, X, t


def test_gradient_matches_finite_differences():
    P, X, t = problem()
    g = cost_mlp.gradient(P, X, t)
    rng = np.random.default_rng(1)
    # every bias and output weight, plus a sample of the weight matrices
    F = X.shape[1]
    idx = list(range(F * 64, F * 64 + 64)) + list(range(cost_mlp.nparams(F) - 65, cost_mlp.nparams(F)))
    idx += list(rng.choice(F * 64, 40, replace=False)) + list(F * 64 + 64 + rng.choice(64 * 64, 60, replace=False))
    h = 1e-6
    for i in idx:
        e = np.zeros_like(P)
        e[i] = h
        fd = (cost_mlp.loss(P + e, X, t) - cost_mlp.loss(P - e, X, t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_forward_matches_loop():
    P, X, t = problem(n=5)
    W1, b1, W2, b2, w3, b3 = cost_mlp.unpack(P, X.shape[1])
    y, _, _ = cost_mlp.forward(P, X)
    for s in range(X.shape[0]):
        h1 = [np.tanh(b1[j] + sum(X[s, f] * W1[f, j] for f in range(X.shape[1]))) for j in range(64)]
        h2 = [np.tanh(b2[j] + sum(h1[i] * W2[i, j] for i in range(64))) for j in range(64)]
        ys = b3 + sum(h2[j] * w3[j] for j in range(64))
        assert abs(ys - y[s]) <= 1e-12 * max(1.0, abs(ys))
