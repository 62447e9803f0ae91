"""Host-side mirror of the reference's XC-build interface over the C ABI.

XCBuilder follows the life cycle of ParaGauss's xc_hamiltonian module:
    xc_setup  -> XCBuilder(basis, grid)            (pgxc_create)
    density   -> set_density(P) / set_orbitals(C, occ)
    xc_build  -> build(functional) -> (ham_xc_arr, exc_int, charge_int)
    xc_close  -> close()
ham_xc_arr is the packed lower triangle (j = 1..n, k = 1..j) of
modules/xc_hamiltonian.f90:1561-1567; exc/charge are exc_int and charge_int
(:1408,1423).  Errors raise PgxcError with the library's message, the role
error_handler plays in the reference.
"""
import ctypes

import numpy as np

from . import _lib

XALPHA, PWLDAC, PBEX, PBEC = 1 << 1, 1 << 3, 1 << 10, 1 << 18
PBE = XALPHA | PWLDAC | PBEX | PBEC
LDA = XALPHA | PWLDAC
FUNCTIONALS = {"PBE": PBE, "LDA": LDA, "XALPHA": XALPHA}

FP64, FP32, BF16X3 = 0, 1, 2
PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16x3": BF16X3}
FULL, INCREMENTAL = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def _ptr(a):
    return a.ctypes.data_as(_dp)


class XCBuilder:
    """tile: function tile of the contractions (64/96/128, 0 = least padding);
    chunk_points: points per work chunk (multiple of 2048, 0 = default);
    variants: (FP64, FP32) contraction kernel variants (-1 = default);
    screen: skip the blocks the radial cut-off makes exactly zero;
    basis_cache: keep every chunk's basis values on the device across builds."""

    def __init__(self, basis, grid, veclen=128, tile=0, chunk_points=0, variants=(-1, -1),
                 screen=True, basis_cache=True):
        self._L = _lib.load()
        self.basis = basis
        self.nbas = basis.nbas
        g = np.asfortranarray(grid, dtype=np.float64)
        if g.ndim != 2 or g.shape[1] != 4:
            raise ValueError("grid must be (npts, 4): x, y, z, w")
        self.npts = g.shape[0]
        self._keep = [np.ascontiguousarray(basis.atom_xyz, dtype=np.float64),
                      np.ascontiguousarray(basis.shell_info, dtype=np.int32),
                      np.ascontiguousarray(basis.exps), np.ascontiguousarray(basis.norms),
                      np.ascontiguousarray(basis.contr), g]
        axyz, info, ex, nm, ct, g = self._keep
        h = ctypes.c_void_p()
        opt = _lib.Options(tile, chunk_points, variants[0], variants[1], 0 if screen else 1,
                           0 if basis_cache else 1)
        rc = self._L.pgxc_create_opt(ctypes.byref(h), basis.n_atoms, _ptr(axyz), basis.n_shells,
                                     info.ctypes.data_as(_ip), _ptr(ex), _ptr(nm), _ptr(ct),
                                     _ptr(g), self.npts, veclen, ctypes.byref(opt))
        self._h = h
        self.ispin = 1
        self._check(rc)

    def _check(self, rc):
        if rc != _lib.PGXC_OK:
            msg = self._L.pgxc_last_error(self._h).decode()
            raise _lib.PgxcError(rc, msg)

    def set_density(self, P, incremental=False):
        """P: (n, n), or (n, n, 2) for an open shell (alpha, beta)."""
        P = np.asfortranarray(P, dtype=np.float64)
        if P.shape[:2] != (self.nbas, self.nbas) or P.ndim not in (2, 3) or \
                (P.ndim == 3 and P.shape[2] != 2):
            raise ValueError("P must be (nbas, nbas) or (nbas, nbas, 2)")
        ispin = 1 if P.ndim == 2 else 2
        self._check(self._L.pgxc_set_density(self._h, ispin, self.nbas, _ptr(P),
                                             INCREMENTAL if incremental else FULL))
        self.ispin = ispin

    def set_orbitals(self, C, occ, incremental=False):
        """C: (n, nocc) and occ (nocc,), or (n, nocc, 2) and (nocc, 2) for an
        open shell."""
        C = np.asfortranarray(C, dtype=np.float64)
        ispin = 1 if C.ndim == 2 else 2
        occ = np.asfortranarray(occ, dtype=np.float64)
        if C.ndim not in (2, 3) or C.shape[0] != self.nbas or occ.shape != C.shape[1:]:
            raise ValueError("C must be (nbas, nocc[, 2]) with occ (nocc[, 2])")
        self._check(self._L.pgxc_set_orbitals(self._h, ispin, self.nbas, C.shape[1], _ptr(C),
                                              _ptr(occ), INCREMENTAL if incremental else FULL))
        self.ispin = ispin

    def build(self, functional=PBE, precision=FP64, fetch=True):
        if isinstance(functional, str):
            functional = FUNCTIONALS[functional]
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        if not fetch:
            self._check(self._L.pgxc_build(self._h, functional, precision, None, None, None))
            return None
        # ispin = 2: alpha / beta interleaved (ham[0::2], ham[1::2])
        ham = np.zeros(self.ispin * self.nbas * (self.nbas + 1) // 2)
        e = np.zeros(1)
        q = np.zeros(1)
        self._check(self._L.pgxc_build(self._h, functional, precision, _ptr(ham), _ptr(e), _ptr(q)))
        return ham, float(e[0]), float(q[0])

    def energy(self, functional=PBE, precision=FP64):
        """Energies of the functional's terms on the current density, without
        the potential (pgxc_energy): ({"xalpha", "pwldac", "pbex", "pbec"}, charge)."""
        if isinstance(functional, str):
            functional = FUNCTIONALS[functional]
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        t = np.zeros(4)
        q = np.zeros(1)
        self._check(self._L.pgxc_energy(self._h, functional, precision, _ptr(t), _ptr(q)))
        return dict(zip(("xalpha", "pwldac", "pbex", "pbec"), t.tolist())), float(q[0])

    def gradient(self):
        """dE_xc/dR (n_atoms, 3), basis-centre term on the fixed grid, at the
        density and potential of the last build (pgxc_gradient)."""
        g = np.zeros((self.basis.n_atoms, 3))
        self._check(self._L.pgxc_gradient(self._h, _ptr(g)))
        return g

    def gradient_grid(self, owner):
        """dE_xc/dR (n_atoms, 3) with the grid moving with the atoms: owner[p]
        is the atom whose atomic grid holds point p (pgxc_gradient_grid)."""
        own = np.ascontiguousarray(owner, dtype=np.int32)
        if own.shape != (self.npts,):
            raise ValueError("owner must have one entry per grid point")
        g = np.zeros((self.basis.n_atoms, 3))
        self._check(self._L.pgxc_gradient_grid(
            self._h, own.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _ptr(g)))
        return g

    def result(self):
        ham = np.zeros(self.ispin * self.nbas * (self.nbas + 1) // 2)
        e = np.zeros(1)
        q = np.zeros(1)
        self._check(self._L.pgxc_get_result(self._h, _ptr(ham), _ptr(e), _ptr(q)))
        return ham, float(e[0]), float(q[0])

    def layout(self):
        out = (ctypes.c_int64 * 8)()
        self._check(self._L.pgxc_layout(self._h, out))
        return dict(NP=out[0], tile=out[1], chunk_points=out[2], chunks=out[3], nsplit=out[4],
                    npad=out[5], variant_fp64=out[6], variant_fp32=out[7])

    def stats(self):
        s = np.zeros(5)
        self._check(self._L.pgxc_last_stats(self._h, _ptr(s)))
        w = np.zeros(3)
        self._check(self._L.pgxc_last_work(self._h, _ptr(w)))
        return dict(points=s[0], batches=s[1], flops=s[2], exps=s[3], bytes=s[4], bn=w[0], n2=w[1])

    def timing(self):
        t = np.zeros(6)
        self._check(self._L.pgxc_last_timing(self._h, _ptr(t)))
        return dict(basis=t[0], density=t[1], functional=t[2], potential=t[3], pack=t[4],
                    total=t[5])


    def set_schedule(self, mode, model=None):
        """mode: _lib.SCHED_STATIC / SCHED_ANALYTIC / SCHED_MODEL (with a CostModel).
        The builder keeps a reference to the model: the library holds its pointer."""
        self._check(self._L.pgxc_set_schedule(self._h, mode, None if model is None else model.handle))
        self._model = model

    def save_state(self):
        """The state incremental builds continue from (pgxc_save_state), as bytes."""
        n = self._L.pgxc_state_bytes(self._h)
        if n < 0:
            self._check(int(-n))
        buf = np.zeros(n, dtype=np.uint8)
        self._check(self._L.pgxc_save_state(self._h, buf.ctypes.data_as(ctypes.c_void_p), n))
        return buf.tobytes()

    def load_state(self, data):
        """Continue from a state saved by save_state on a builder of the same
        geometry (pgxc_load_state)."""
        buf = np.frombuffer(data, dtype=np.uint8)
        self._check(self._L.pgxc_load_state(self._h, buf.ctypes.data_as(ctypes.c_void_p), buf.size))
        # the record's spin count: int32 after the 8-byte magic, the 8-byte
        # geometry hash, five int64 sizes and the uint32 functional mask
        self.ispin = int(np.frombuffer(data, dtype=np.int32, count=1, offset=60)[0])

    def set_lanes(self, lanes):
        """1: every chunk on one stream; 2 (default for grids of several chunks):
        chunks alternate between two streams (pgxc_set_lanes)."""
        self._check(self._L.pgxc_set_lanes(self._h, int(lanes)))

    def set_item_log(self, on=True):
        self._check(self._L.pgxc_set_item_log(self._h, 1 if on else 0))

    def item_log(self):
        """(features (n, NFEATURES), microseconds (n)) of the last logged build."""
        n = self._L.pgxc_item_log(self._h, None, None, 0)
        if n < 0:
            raise _lib.PgxcError(-n, "item log")
        F = np.zeros((n, _lib.NFEATURES))
        us = np.zeros(n)
        m = self._L.pgxc_item_log(self._h, _ptr(F), _ptr(us), n)
        if m != n:
            raise _lib.PgxcError(3, "item log")
        return F, us

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.pgxc_destroy(self._h)
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


class Batch:
    """Many molecules built together (pgxc_batch_*): one launch per stage over
    the work items of all of them, optionally replayed from a hipGraph.
    Results stay in each XCBuilder (builder.result())."""

    def __init__(self, builders):
        self._L = _lib.load()
        self.builders = list(builders)
        arr = (ctypes.c_void_p * len(self.builders))(*[b._h.value for b in self.builders])
        h = ctypes.c_void_p()
        rc = self._L.pgxc_batch_create(ctypes.byref(h), len(self.builders), arr)
        self._h = h
        self._check(rc)

    def _check(self, rc):
        if rc != _lib.PGXC_OK:
            raise _lib.PgxcError(rc, self._L.pgxc_batch_last_error(self._h).decode())

    def set_priority(self, cost):
        if cost is None:
            self._check(self._L.pgxc_batch_set_priority(self._h, None))
        else:
            c = np.ascontiguousarray(cost, dtype=np.float64)
            assert c.size == len(self.builders)
            self._check(self._L.pgxc_batch_set_priority(self._h, _ptr(c)))

    def build(self, functional=PBE, precision=FP64, graph=False):
        if isinstance(functional, str):
            functional = FUNCTIONALS[functional]
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        self._check(self._L.pgxc_batch_build(self._h, functional, precision, int(bool(graph))))

    def last_ms(self):
        t = np.zeros(1)
        self._check(self._L.pgxc_batch_last_ms(self._h, _ptr(t)))
        return float(t[0])

    def results(self):
        return [b.result() for b in self.builders]


    def set_schedule(self, mode, model=None):
        """mode: _lib.SCHED_STATIC / SCHED_ANALYTIC / SCHED_MODEL (with a CostModel).
        The batch keeps a reference to the model: the library holds its pointer."""
        self._check(self._L.pgxc_batch_set_schedule(self._h, mode, None if model is None else model.handle))
        self._model = model

    def set_item_log(self, on=True):
        self._check(self._L.pgxc_batch_set_item_log(self._h, 1 if on else 0))

    def item_log(self):
        """(features (n, NFEATURES), microseconds (n)) of the last logged build."""
        n = self._L.pgxc_batch_item_log(self._h, None, None, 0)
        if n < 0:
            raise _lib.PgxcError(-n, "item log")
        F = np.zeros((n, _lib.NFEATURES))
        us = np.zeros(n)
        m = self._L.pgxc_batch_item_log(self._h, _ptr(F), _ptr(us), n)
        if m != n:
            raise _lib.PgxcError(3, "item log")
        return F, us

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.pgxc_batch_destroy(self._h)
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


def unpack(ham_packed, n):
    """Packed lower triangle -> symmetric n x n."""
    V = np.zeros((n, n))
    j, k = np.tril_indices(n)
    V[j, k] = ham_packed
    V[k, j] = ham_packed
    return V


def becke_weights(atom_xyz, pts, owner, w_atomic):
    """Becke-partitioned weights on the GPU (pgxc_becke_weights)."""
    L = _lib.load()
    axyz = np.ascontiguousarray(atom_xyz, dtype=np.float64)
    P = np.asfortranarray(pts, dtype=np.float64)
    own = np.ascontiguousarray(owner, dtype=np.int32)
    wa = np.ascontiguousarray(w_atomic, dtype=np.float64)
    out = np.zeros(P.shape[0])
    rc = L.pgxc_becke_weights(axyz.shape[0], _ptr(axyz), P.shape[0], _ptr(P),
                              own.ctypes.data_as(_ip), _ptr(wa), _ptr(out))
    if rc != 0:
        raise _lib.PgxcError(rc, "pgxc_becke_weights failed")
    return out
