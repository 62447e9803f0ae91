"""Learned cost model of the contraction work items (pgxc_costmodel_*).

A 3-layer MLP (NFEATURES -> 64 -> 64 -> 1, tanh) predicting log(item time in
us) from the item features the library logs (pgxc_item_log).  Training,
gradients and inference run on the GPU in the library's own HIP kernels.
The schedules that use it (SCHED_MODEL) replace the static contiguous split
of the reference's distribute_jobs (dlb/dlb_common.f90:516-561).
"""
import ctypes
import os

import numpy as np

from . import _lib

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


class CostModel:
    def __init__(self, n_features=_lib.NFEATURES, seed=0):
        self._L = _lib.load()
        h = ctypes.c_void_p()
        rc = self._L.pgxc_costmodel_create(ctypes.byref(h), n_features, seed)
        self._h = h
        self.n_features = n_features
        self._check(rc)

    @property
    def handle(self):
        return self._h

    def _check(self, rc):
        if rc != _lib.PGXC_OK:
            raise _lib.PgxcError(rc, self._L.pgxc_costmodel_last_error(self._h).decode())

    def _xy(self, X, y=None):
        X = np.ascontiguousarray(X, dtype=np.float64)
        assert X.ndim == 2 and X.shape[1] == self.n_features
        if y is not None:
            y = np.ascontiguousarray(y, dtype=np.float64)
            assert y.shape == (X.shape[0],)
        return X, y

    @property
    def nparams(self):
        return int(self._L.pgxc_costmodel_nparams(self._h))

    def params(self):
        p = np.zeros(self.nparams)
        nrm = np.zeros(2 * self.n_features + 2)
        self._check(self._L.pgxc_costmodel_params(self._h, _ptr(p), _ptr(nrm)))
        return p, nrm

    def set_params(self, p, norm=None):
        p = np.ascontiguousarray(p, dtype=np.float64)
        assert p.size == self.nparams
        nrm = None if norm is None else np.ascontiguousarray(norm, dtype=np.float64)
        self._check(self._L.pgxc_costmodel_set_params(self._h, _ptr(p),
                                                      None if nrm is None else _ptr(nrm)))

    def train(self, X, y, steps, lr=1e-2):
        X, y = self._xy(X, y)
        loss = np.zeros(1)
        self._check(self._L.pgxc_costmodel_train(self._h, _ptr(X), _ptr(y), X.shape[0], steps, lr,
                                                 _ptr(loss)))
        return float(loss[0])

    def gradient(self, X, y):
        X, y = self._xy(X, y)
        g = np.zeros(self.nparams)
        loss = np.zeros(1)
        self._check(self._L.pgxc_costmodel_gradient(self._h, _ptr(X), _ptr(y), X.shape[0], _ptr(g),
                                                    _ptr(loss)))
        return g, float(loss[0])

    def predict(self, X):
        X, _ = self._xy(X)
        out = np.zeros(X.shape[0])
        self._check(self._L.pgxc_costmodel_predict(self._h, _ptr(X), X.shape[0], _ptr(out)))
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.pgxc_costmodel_destroy(self._h)
        self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def target(us):
    """Training target of an item time: log(us), with a floor of 0.01 us."""
    return np.log(np.maximum(np.asarray(us, dtype=np.float64), 1e-2))


DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cost_model_mi355x.npz")


def default_model(path=DEFAULT_PATH):
    """The cost model shipped with the package (trained on MI355X item timings
    of the H2O and (H2O)32 builds by `bench.py --workload learned
    --save-model`), or None if the file is absent."""
    if not os.path.exists(path):
        return None
    d = np.load(path, allow_pickle=False)
    m = CostModel(_lib.NFEATURES, seed=0)
    m.set_params(d["params"], d["norm"])
    return m
