"""Loader of libpgxc.so, the HIP implementation behind include/pgxc.h.

There is no CPU fallback: if the library is missing, or no HIP device is
present when a build is requested, the calls raise.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libpgxc.so")

PGXC_OK = 0
ERRORS = {1: "PGXC_ERR_ARG", 2: "PGXC_ERR_UNSUPPORTED", 3: "PGXC_ERR_HIP",
          4: "PGXC_ERR_NOMEM", 5: "PGXC_ERR_STATE"}

# header symbols, checked by the tests against include/pgxc.h
EXPORTS = ("pgxc_create", "pgxc_create_opt", "pgxc_layout", "pgxc_set_density",
           "pgxc_set_orbitals", "pgxc_build", "pgxc_get_result", "pgxc_energy", "pgxc_gradient", "pgxc_gradient_grid",
           "pgxc_nbas",
           "pgxc_last_stats", "pgxc_last_work",
           "pgxc_last_timing", "pgxc_last_error", "pgxc_destroy", "pgxc_becke_weights",
           "pgxc_becke_weights_noalloc",
           "pgxc_batch_create", "pgxc_batch_set_priority", "pgxc_batch_build",
           "pgxc_batch_last_ms", "pgxc_batch_last_error", "pgxc_batch_destroy",
           "pgxc_set_schedule", "pgxc_batch_set_schedule", "pgxc_set_lanes", "pgxc_set_item_log",
           "pgxc_state_bytes", "pgxc_save_state", "pgxc_load_state",
           "pgxc_batch_set_item_log", "pgxc_item_log", "pgxc_batch_item_log",
           "pgxc_costmodel_create", "pgxc_costmodel_nparams", "pgxc_costmodel_params",
           "pgxc_costmodel_set_params", "pgxc_costmodel_train", "pgxc_costmodel_gradient",
           "pgxc_costmodel_predict", "pgxc_costmodel_last_error", "pgxc_costmodel_destroy")

NFEATURES = 9
SCHED_STATIC, SCHED_ANALYTIC, SCHED_MODEL = 0, 1, 2


class Options(ctypes.Structure):
    _fields_ = [("tile", ctypes.c_int32), ("chunk_points", ctypes.c_int64),
                ("variant_fp64", ctypes.c_int32), ("variant_fp32", ctypes.c_int32),
                ("no_screen", ctypes.c_int32), ("no_basis_cache", ctypes.c_int32)]

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p
_lib = None


class PgxcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERRORS.get(code, "status %d" % code), msg))
        self.code = code


def load():
    """Return the ctypes handle; raise if libpgxc.so has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libpgxc.so not built (expected at %s); run "
                           "__graft_entry__.build() or make -C paragauss-gpl_amd" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    L.pgxc_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int32, _dp, ctypes.c_int32, _ip,
                              _dp, _dp, _dp, _dp, ctypes.c_int64, ctypes.c_int32]
    L.pgxc_create_opt.argtypes = L.pgxc_create.argtypes + [ctypes.POINTER(Options)]
    L.pgxc_layout.argtypes = [_vp, ctypes.POINTER(ctypes.c_int64)]
    L.pgxc_batch_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int32, ctypes.POINTER(_vp)]
    L.pgxc_batch_set_priority.argtypes = [_vp, _dp]
    L.pgxc_batch_build.argtypes = [_vp, ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32]
    L.pgxc_batch_last_ms.argtypes = [_vp, _dp]
    L.pgxc_batch_last_error.argtypes = [_vp]
    L.pgxc_batch_last_error.restype = ctypes.c_char_p
    L.pgxc_batch_destroy.argtypes = [_vp]
    L.pgxc_set_density.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, _dp, ctypes.c_int32]
    L.pgxc_set_orbitals.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _dp,
                                    _dp, ctypes.c_int32]
    L.pgxc_build.argtypes = [_vp, ctypes.c_uint32, ctypes.c_int32, _dp, _dp, _dp]
    L.pgxc_get_result.argtypes = [_vp, _dp, _dp, _dp]
    L.pgxc_energy.argtypes = [_vp, ctypes.c_uint32, ctypes.c_int32, _dp, _dp]
    L.pgxc_gradient.argtypes = [_vp, _dp]
    L.pgxc_gradient_grid.argtypes = [_vp, ctypes.POINTER(ctypes.c_int32), _dp]
    L.pgxc_nbas.argtypes = [_vp]
    L.pgxc_nbas.restype = ctypes.c_int64
    L.pgxc_last_stats.argtypes = [_vp, _dp]
    L.pgxc_last_work.argtypes = [_vp, _dp]
    L.pgxc_last_timing.argtypes = [_vp, _dp]
    L.pgxc_last_error.argtypes = [_vp]
    L.pgxc_last_error.restype = ctypes.c_char_p
    L.pgxc_destroy.argtypes = [_vp]
    L.pgxc_becke_weights.argtypes = [ctypes.c_int32, _dp, ctypes.c_int64, _dp, _ip, _dp, _dp]
    L.pgxc_becke_weights.restype = ctypes.c_int
    L.pgxc_becke_weights_noalloc.argtypes = [ctypes.c_int32, _vp, ctypes.c_int64, _vp, _vp, _vp,
                                             _vp, _vp]
    L.pgxc_becke_weights_noalloc.restype = ctypes.c_int
    L.pgxc_set_schedule.argtypes = [_vp, ctypes.c_int32, _vp]
    L.pgxc_batch_set_schedule.argtypes = [_vp, ctypes.c_int32, _vp]
    L.pgxc_set_lanes.argtypes = [_vp, ctypes.c_int32]
    L.pgxc_state_bytes.argtypes = [_vp]
    L.pgxc_save_state.argtypes = [_vp, _vp, ctypes.c_int64]
    L.pgxc_load_state.argtypes = [_vp, _vp, ctypes.c_int64]
    L.pgxc_set_item_log.argtypes = [_vp, ctypes.c_int32]
    L.pgxc_batch_set_item_log.argtypes = [_vp, ctypes.c_int32]
    L.pgxc_item_log.argtypes = [_vp, _dp, _dp, ctypes.c_int64]
    L.pgxc_batch_item_log.argtypes = [_vp, _dp, _dp, ctypes.c_int64]
    L.pgxc_costmodel_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int32, ctypes.c_uint64]
    L.pgxc_costmodel_nparams.argtypes = [_vp]
    L.pgxc_costmodel_params.argtypes = [_vp, _dp, _dp]
    L.pgxc_costmodel_set_params.argtypes = [_vp, _dp, _dp]
    L.pgxc_costmodel_train.argtypes = [_vp, _dp, _dp, ctypes.c_int64, ctypes.c_int32,
                                       ctypes.c_double, _dp]
    L.pgxc_costmodel_gradient.argtypes = [_vp, _dp, _dp, ctypes.c_int64, _dp, _dp]
    L.pgxc_costmodel_predict.argtypes = [_vp, _dp, ctypes.c_int64, _dp]
    L.pgxc_costmodel_last_error.argtypes = [_vp]
    L.pgxc_costmodel_last_error.restype = ctypes.c_char_p
    L.pgxc_costmodel_destroy.argtypes = [_vp]
    special = {"pgxc_nbas": ctypes.c_int64, "pgxc_last_error": ctypes.c_char_p,
               "pgxc_batch_last_error": ctypes.c_char_p,
               "pgxc_costmodel_last_error": ctypes.c_char_p,
               "pgxc_costmodel_nparams": ctypes.c_int32, "pgxc_item_log": ctypes.c_int64,
               "pgxc_batch_item_log": ctypes.c_int64, "pgxc_state_bytes": ctypes.c_int64}
    for name in EXPORTS:
        getattr(L, name).restype = special.get(name, ctypes.c_int)
    _lib = L
    return L
