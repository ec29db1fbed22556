"""ctypes view of include/photonprop.h and loader for libphotonprop.so.

The library is the product: if it is missing this module raises instead of
falling back to anything else.
"""
import ctypes
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libphotonprop.so")

PP_OK = 0
PP_ERR_INVALID_ARG = -1
PP_ERR_OVERFLOW = -2
PP_ERR_DEVICE = -3
PP_ERR_OOM = -4
PP_ERR_UNSUPPORTED = -5
PP_ERR_BUSY = -6
PP_ERR_NOT_FOUND = -7

PP_SPEC_MONO = 0
PP_SPEC_KNOTS = 1
PP_SPEC_INV_LAMBDA2 = 2

PP_EMIT_CHERENKOV_CONE = 0
PP_EMIT_ISOTROPIC = 1
PP_EMIT_REF_ISOTROPIC = 2

PP_MLP_SILU = 0
PP_MLP_RELU = 1

PP_MAX_TABLE_POINTS = 2048
PP_MAX_KNOTS = 1024
PP_MAX_STEPS = 1 << 24

c_float_p = ctypes.POINTER(ctypes.c_float)


class pp_medium_tables(ctypes.Structure):
    _fields_ = [
        ("lambda_min", ctypes.c_float),
        ("lambda_max", ctypes.c_float),
        ("n_points", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("n_phase", c_float_p),
        ("abs_len", c_float_p),
        ("sca_len", c_float_p),
        ("g", ctypes.c_float),
        ("reserved2", ctypes.c_float),
    ]


class pp_spectrum(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_uint32),
        ("n_knots", ctypes.c_uint32),
        ("lo", ctypes.c_float),
        ("hi", ctypes.c_float),
        ("mono", ctypes.c_float),
        ("reserved", ctypes.c_float),
        ("knots", c_float_p),
    ]


class pp_emitter(ctypes.Structure):
    _fields_ = [
        ("pos", ctypes.c_float * 3),
        ("t0", ctypes.c_float),
        ("axis", ctypes.c_float * 3),
        ("beta", ctypes.c_float),
        ("n_photons", ctypes.c_uint64),
        ("profile", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class pp_sphere(ctypes.Structure):
    _fields_ = [("center", ctypes.c_float * 3), ("radius", ctypes.c_float)]


class pp_summary(ctypes.Structure):
    _fields_ = [
        ("n_sim", ctypes.c_uint64),
        ("n_steps", ctypes.c_uint64),
        ("n_hits", ctypes.c_uint64),
        ("n_stored", ctypes.c_uint64),
        ("sum_w", ctypes.c_double),
        ("sum_wt", ctypes.c_double),
        ("sum_wt2", ctypes.c_double),
        ("status", ctypes.c_int32),
        ("reserved", ctypes.c_uint32),
    ]


class pp_request(ctypes.Structure):
    _fields_ = [
        ("emitter", pp_emitter),
        ("target", pp_sphere),
        ("seed", ctypes.c_uint64),
        ("stream", ctypes.c_uint32),
        ("max_steps", ctypes.c_uint32),
    ]


PP_SERVER_HOST_POOL = 1
PP_SERVER_LEAVE_ROOM = 2


class pp_server_config(ctypes.Structure):
    _fields_ = [
        ("max_slots", ctypes.c_uint32),
        ("max_tiles_per_request", ctypes.c_uint32),
        ("hit_pool_records", ctypes.c_uint64),
        ("idle_exit_us", ctypes.c_uint32),
        ("device_wait_us", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class pp_server_stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "submitted", "completed", "in_flight", "pending", "kernel_launches",
        "pool_records_free", "tiles_published", "grid_blocks", "device_stops",
        "kernel_running", "host_pool")]


PP_TILE_PHOTONS = 8192
PP_MAX_HIT_RECORDS = 1 << 36
PP_NOALLOC_MAX_REQUESTS = 4096
PP_NOALLOC_MAX_TILES = 1 << 20
PP_NOALLOC_MAX_HITS = 1 << 18

# pp_hit as a numpy record (64 bytes).
HIT_DTYPE = np.dtype([
    ("pos", "<f4", (3,)),
    ("time", "<f4"),
    ("dir", "<f4", (3,)),
    ("wavelength", "<f4"),
    ("init_dir", "<f4", (3,)),
    ("weight", "<f4"),
    ("dist", "<f4"),
    ("request", "<u4"),
    ("photon_id", "<u8"),
])
assert HIT_DTYPE.itemsize == 64

SUMMARY_DTYPE = np.dtype([
    ("n_sim", "<u8"), ("n_steps", "<u8"), ("n_hits", "<u8"), ("n_stored", "<u8"),
    ("sum_w", "<f8"), ("sum_wt", "<f8"), ("sum_wt2", "<f8"),
    ("status", "<i4"), ("reserved", "<u4"),
])
assert SUMMARY_DTYPE.itemsize == ctypes.sizeof(pp_summary) == 64

# Every symbol include/photonprop.h declares, with (restype, argtypes).
_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32
SIGNATURES = {
    "pp_last_error": (ctypes.c_char_p, []),
    "pp_abi_version": (ctypes.c_int, []),
    "pp_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(pp_medium_tables)]),
    "pp_set_medium": (ctypes.c_int, [_vp, ctypes.POINTER(pp_medium_tables)]),
    "pp_destroy": (ctypes.c_int, [_vp]),
    "pp_device_info": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int)]),
    "pp_propagate": (ctypes.c_int, [_vp, ctypes.POINTER(pp_emitter), ctypes.POINTER(pp_spectrum),
                                    ctypes.POINTER(pp_sphere), _u64, _u32, _u32, _vp, _u64,
                                    ctypes.POINTER(pp_summary)]),
    "pp_propagate_batch": (ctypes.c_int, [_vp, ctypes.POINTER(pp_request), _u32,
                                          ctypes.POINTER(pp_spectrum), _vp, _u64, _u64, _vp, _vp]),
    "pp_propagate_batch_device": (ctypes.c_int, [_vp, ctypes.POINTER(pp_request), _u32,
                                                 ctypes.POINTER(pp_spectrum), _vp, _u64, _u64, _vp,
                                                 _vp]),
    "pp_propagate_batch_device_noalloc": (ctypes.c_int, [_vp, ctypes.POINTER(pp_request), _u32,
                                                         ctypes.POINTER(pp_spectrum), _vp, _u64, _u64,
                                                         _vp, _vp]),
    "pp_synchronize": (ctypes.c_int, [_vp]),
    "pp_last_kernel_ms": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
    "pp_server_create": (ctypes.c_int, [_vp, ctypes.POINTER(pp_spectrum), ctypes.POINTER(pp_server_config),
                                        ctypes.POINTER(_vp)]),
    "pp_server_submit": (ctypes.c_int, [_vp, ctypes.POINTER(pp_request), _u64, ctypes.POINTER(_u64)]),
    "pp_server_wait": (ctypes.c_int, [_vp, _u64, _u32, ctypes.POINTER(ctypes.c_int)]),
    "pp_server_result": (ctypes.c_int, [_vp, _u64, ctypes.POINTER(pp_summary), _vp, _u64,
                                        ctypes.POINTER(_u64), ctypes.POINTER(ctypes.c_double)]),
    "pp_server_stats_get": (ctypes.c_int, [_vp, ctypes.POINTER(pp_server_stats)]),
    "pp_server_destroy": (ctypes.c_int, [_vp]),
    "pp_mlp_create": (ctypes.c_int, [_vp, _u32, _u32, _u32, _u32, _u64, ctypes.POINTER(_vp)]),
    "pp_mlp_destroy": (ctypes.c_int, [_vp]),
    "pp_mlp_n_params": (ctypes.c_int, [_vp, ctypes.POINTER(_u64)]),
    "pp_mlp_get_params": (ctypes.c_int, [_vp, _vp, _u64]),
    "pp_mlp_set_params": (ctypes.c_int, [_vp, _vp, _u64]),
    "pp_mlp_set_activation": (ctypes.c_int, [_vp, _u32]),
    "pp_mlp_forward_device": (ctypes.c_int, [_vp, _vp, _u64, _vp]),
    "pp_mlp_forward": (ctypes.c_int, [_vp, _vp, _u64, _vp]),
    "pp_mlp_last_forward_ms": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
    "pp_mlp_loss_grad_device": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _u32, _u32, ctypes.c_float,
                                               _vp, _vp]),
    "pp_mlp_train_step_device": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _u32, _u32, ctypes.c_float,
                                                ctypes.c_float, _vp]),
    "pp_mlp_train_steps_device": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _u32, _u32, _u32, _vp,
                                                 ctypes.c_float, _vp]),
    "pp_selftest_rotate": (ctypes.c_int, [_vp, _vp, _vp, _vp, _u64]),
    "pp_selftest_hg": (ctypes.c_int, [_vp, ctypes.c_float, _u64, _vp, _u64]),
    "pp_selftest_philox": (ctypes.c_int, [_vp, _vp, _u64, _vp, _u64]),
    "pp_selftest_hg_moments": (ctypes.c_int, [_vp, ctypes.c_float, _u64, _u64, _vp]),
}

_lib = None


class PhotonPropError(RuntimeError):
    """A library call returned a negative status."""

    def __init__(self, code, message):
        super().__init__("%s (status %d)" % (message, code))
        self.code = code


def load_library(path=None):
    """Load libphotonprop.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or os.environ.get("PHOTONPROP_LIB", LIB_PATH)
    if not os.path.exists(p):
        raise OSError("libphotonprop.so not found at %s; run __graft_entry__.build()" % p)
    # One HIP runtime per process.  PyTorch ships its own libamdhip64 (SONAME
    # libamdhip64.so.7) and links it as "libamdhip64.so"; loaded first, it also
    # satisfies this library's libamdhip64.so.7.  Loaded after the system
    # runtime, it would come in as a second runtime that sees no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc, lib=None):
    """Raise PhotonPropError for a negative status (unsupported kinds raise the
    ArgumentError-like ValueError the reference's dispatch fallback throws,
    src/photon_prop_cuda.jl:67, :70)."""
    if rc >= 0:
        return rc
    lib = lib or load_library()
    msg = lib.pp_last_error().decode(errors="replace")
    if rc == PP_ERR_UNSUPPORTED:
        raise ValueError(msg)
    raise PhotonPropError(rc, msg)


def ptr(a):
    """Data pointer of a numpy array (or None for an empty one)."""
    if a is None:
        return None
    return ctypes.c_void_p(a.ctypes.data) if a.size else None
