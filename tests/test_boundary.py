"""The C ABI (include/pgxc.h) without a GPU: the library loads, exports every
declared entry point, and rejects invalid input with the documented status
before touching a device.  The Fortran binding declares the same names."""
import ctypes
import os
import re

import numpy as np
import pytest

from pgxc import _lib, basis, systems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pgxc.h")
FBIND = os.path.join(ROOT, "paragauss-gpl_amd", "fortran", "pgxc_binding.f90")


def declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int32_t|int64_t|const char \*)\s*(pgxc_\w+)\s*\(", src, re.M)))


def test_header_declares_expected_api():
    names = declared()
    assert set(names) == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    for name in declared():
        assert hasattr(L, name), name


def test_fortran_binding_names_every_symbol():
    src = open(FBIND).read()
    for name in declared():
        assert 'bind(C, name="%s")' % name in src, name


def test_fortran_driver_built():
    exe = os.path.join(ROOT, "paragauss-gpl_amd", "fortran", "pgxc_driver")
    assert os.path.exists(exe) and os.access(exe, os.X_OK)


def _create(b, grid, veclen=128):
    L = _lib.load()
    h = ctypes.c_void_p()
    dp = ctypes.POINTER(ctypes.c_double)
    axyz = np.ascontiguousarray(b.atom_xyz)
    info = np.ascontiguousarray(b.shell_info, dtype=np.int32)
    ex, nm, ct = b.exps, b.norms, np.ascontiguousarray(b.contr)
    g = np.asfortranarray(grid)
    rc = L.pgxc_create(ctypes.byref(h), b.n_atoms, axyz.ctypes.data_as(dp), b.n_shells,
                       info.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ex.ctypes.data_as(dp),
                       nm.ctypes.data_as(dp), ct.ctypes.data_as(dp), g.ctypes.data_as(dp),
                       g.shape[0], veclen)
    msg = L.pgxc_last_error(h).decode() if h.value else ""
    L.pgxc_destroy(h)
    return rc, msg


def test_create_rejects_high_l_before_device_use():
    b = basis.Basis(np.zeros((1, 3)), [basis.Shell(0, 5, [1.0], 1, np.zeros((1, 0)))])
    rc, msg = _create(b, np.zeros((4, 4)))
    assert rc == 2 and "l = 5" in msg


def test_create_rejects_non_contiguous_atoms():
    sh = [basis.Shell(0, 0, [1.0], 1, np.zeros((1, 0))), basis.Shell(1, 0, [1.0], 1, np.zeros((1, 0))),
          basis.Shell(0, 1, [1.0], 1, np.zeros((1, 0)))]
    b = basis.Basis(np.zeros((2, 3)), sh)
    rc, msg = _create(b, np.zeros((4, 4)))
    assert rc == 1 and "contiguous" in msg


def test_create_rejects_bad_counts_and_exponents():
    b = basis.Basis(np.zeros((1, 3)), [basis.Shell(0, 0, [1.0, 2.0], 3, np.zeros((2, 0)))])
    assert _create(b, np.zeros((4, 4)))[0] == 1
    b = basis.Basis(np.zeros((1, 3)), [basis.Shell(0, 0, [-1.0], 1, np.zeros((1, 0)))])
    assert _create(b, np.zeros((4, 4)))[0] == 1
    b = basis.Basis(np.zeros((1, 3)), [basis.Shell(0, 0, [1.0], 0, np.ones((1, 9)))])
    assert _create(b, np.zeros((4, 4)))[0] == 2  # more than 8 contractions


def test_create_rejects_empty_grid():
    els, xyz = systems.h2o()
    b = basis.seeded_basis(els, xyz, seed=1)
    assert _create(b, np.zeros((0, 4)))[0] == 1


def test_becke_rejects_bad_owner():
    L = _lib.load()
    dp = ctypes.POINTER(ctypes.c_double)
    axyz = np.zeros((2, 3))
    axyz[1, 0] = 1.0
    pts = np.zeros((3, 3), order="F")
    own = np.array([0, 1, 2], dtype=np.int32)
    w = np.ones(3)
    out = np.zeros(3)
    rc = L.pgxc_becke_weights(2, axyz.ctypes.data_as(dp), 3, pts.ctypes.data_as(dp),
                              own.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                              w.ctypes.data_as(dp), out.ctypes.data_as(dp))
    assert rc == 1


def test_null_context_calls_fail_cleanly():
    L = _lib.load()
    assert L.pgxc_build(None, 1 << 1, 0, None, None, None) == 1
    assert L.pgxc_nbas(None) == -1
    assert L.pgxc_destroy(None) == 0
    assert L.pgxc_last_error(None) == b"null context"


def test_product_has_no_cpu_fallback():
    # the product package never imports the oracle
    pkg = os.path.join(ROOT, "paragauss-gpl_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h", ".f90")):
                src = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in src and "liboracle" not in src, f
