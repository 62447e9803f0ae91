"""The allocation-free Becke weights entry point (pgxc_becke_weights_noalloc)
against the allocating one (pgxc_becke_weights), through the C ABI: device
inputs and outputs owned by the caller, the inverse-distance table in the
library's file-scope scratch array.  Two calls in a row, the second on a
smaller input, each equal bit for bit to the allocating entry point; a
size beyond the scratch array's maximum is refused."""
import ctypes

import numpy as np
import pytest

from pgxc import _lib, grid, systems

pytestmark = pytest.mark.gpu

_vp = ctypes.c_void_p


def hip():
    """The HIP runtime the library links (already loaded with it)."""
    _lib.load()
    H = ctypes.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t]
    H.hipFree.argtypes = [_vp]
    H.hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int]
    H.hipDeviceSynchronize.argtypes = []
    return H


H2D, D2H = 1, 2


class Dev:
    """Device copies of host arrays (hipMalloc of the test, not of the library)."""

    def __init__(self, H):
        self.H, self.bufs = H, []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = _vp()
        assert self.H.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.H.hipMemcpy(p, a.ctypes.data_as(_vp), a.nbytes, H2D) == 0
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype=dtype)
        assert self.H.hipMemcpy(out.ctypes.data_as(_vp), p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.H.hipFree(p)


def case(n_water, nrad, seed):
    els, xyz = systems.water_cluster(n_water, seed=seed)
    g = grid.molecular_grid(els, xyz, nrad=nrad, nang=302, partition="none")
    own = grid.owners(els, nrad, 302)
    assert len(own) == g.shape[0]
    return xyz, g, own


def test_noalloc_becke_equals_allocating_entry_point():
    L = _lib.load()
    H = hip()
    dp = ctypes.POINTER(ctypes.c_double)
    for n_water, nrad, seed in ((8, 40, 3), (3, 20, 4)):  # the second call smaller
        xyz, g, own = case(n_water, nrad, seed)
        npts, na = g.shape[0], xyz.shape[0]
        print("noalloc Becke: %d atoms, %d points" % (na, npts), flush=True)
        axyz = np.ascontiguousarray(xyz, dtype=np.float64)
        pts = np.asfortranarray(g[:, :3])
        wat = np.ascontiguousarray(g[:, 3])
        own32 = np.ascontiguousarray(own, dtype=np.int32)
        ref = np.zeros(npts)
        rc = L.pgxc_becke_weights(na, axyz.ctypes.data_as(dp), npts, pts.ctypes.data_as(dp),
                                  own32.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                  wat.ctypes.data_as(dp), ref.ctypes.data_as(dp))
        assert rc == 0
        D = Dev(H)
        try:
            d_a = D.put(axyz)
            d_p = D.put(np.ascontiguousarray(pts.T.reshape(-1)))  # npts x 3 column-major
            d_o = D.put(own32)
            d_w = D.put(wat)
            d_out = D.put(np.full(npts, np.nan))
            rc = L.pgxc_becke_weights_noalloc(na, d_a, npts, d_p, d_o, d_w, d_out, None)
            assert rc == 0, L.pgxc_last_error(None)
            assert H.hipDeviceSynchronize() == 0
            got = D.get(d_out, npts, np.float64)
        finally:
            D.free()
        assert np.array_equal(got, ref)
        assert np.all(np.isfinite(got)) and np.all(got >= 0.0)
    # beyond the scratch array's 1024 atoms: refused before any launch
    rc = L.pgxc_becke_weights_noalloc(1025, _vp(8), 10, _vp(8), _vp(8), _vp(8), _vp(8), None)
    assert rc == 1  # PGXC_ERR_ARG
    assert b"1024" in L.pgxc_last_error(None)
