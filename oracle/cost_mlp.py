"""CPU FP64 restatement of the cost model's network (numpy).

TEST INFRASTRUCTURE ONLY: the checker of the library's HIP forward, backward
and training kernels (paragauss-gpl_amd/csrc/cost_model.hip).  The reference
has no cost model (its batches are split statically, dlb/dlb_common.f90:
516-561); this restates the textbook MLP and its back-propagation, and the
gradient is pinned by central finite differences in the tests.

Network: F -> H -> H -> 1, tanh hidden units, parameters flat as
W1[F][H], b1[H], W2[H][H], b2[H], w3[H], b3.  Loss: mean squared error.
"""
import numpy as np

H = 64


def nparams(F, h=H):
    return F * h + h + h * h + h + h + 1


def unpack(P, F, h=H):
    P = np.asarray(P, dtype=np.float64)
    o = 0
    W1 = P[o:o + F * h].reshape(F, h); o += F * h
    b1 = P[o:o + h]; o += h
    W2 = P[o:o + h * h].reshape(h, h); o += h * h
    b2 = P[o:o + h]; o += h
    w3 = P[o:o + h]; o += h
    b3 = P[o]
    return W1, b1, W2, b2, w3, b3


def forward(P, X):
    X = np.asarray(X, dtype=np.float64)
    W1, b1, W2, b2, w3, b3 = unpack(P, X.shape[1])
    H1 = np.tanh(X @ W1 + b1)
    H2 = np.tanh(H1 @ W2 + b2)
    return H2 @ w3 + b3, H1, H2


def loss(P, X, t):
    y, _, _ = forward(P, X)
    return float(np.mean((y - t) ** 2))


def gradient(P, X, t):
    """d loss / d P by back-propagation, in the flat parameter order."""
    X = np.asarray(X, dtype=np.float64)
    n, F = X.shape
    W1, b1, W2, b2, w3, b3 = unpack(P, F)
    y, H1, H2 = forward(P, X)
    g = 2.0 * (y - t) / n                  # dL/dy
    dA2 = np.outer(g, w3) * (1.0 - H2 ** 2)
    dA1 = (dA2 @ W2.T) * (1.0 - H1 ** 2)
    parts = [X.T @ dA1, dA1.sum(0), H1.T @ dA2, dA2.sum(0), H2.T @ g, np.array([g.sum()])]
    return np.concatenate([p.ravel() for p in parts])
