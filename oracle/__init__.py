"""CPU oracle for the photon-propagation path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
this module, and only as the checker or the timed CPU comparison.  The product
package never imports it.

Python binding of liboracle.so (pp_oracle.c, a C restatement of
src/photon_prop_cuda.jl:301-380 and helpers; see that file's header for the
per-function citations, deviations D1-D10 and how each function is pinned).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-B", "-C", HERE, "liboracle.so"])
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        for suf, real in (("_f32", ctypes.c_float), ("_f64", ctypes.c_double)):
            getattr(L, "or_hg_costheta" + suf).restype = real
            getattr(L, "or_hg_costheta" + suf).argtypes = [real, real]
            getattr(L, "or_rotate_to_axis" + suf).argtypes = [vp, vp, vp]
            getattr(L, "or_update_position" + suf).argtypes = [vp, vp, real, vp]
            getattr(L, "or_ray_sphere" + suf).restype = ctypes.c_int
            getattr(L, "or_ray_sphere" + suf).argtypes = [vp, vp, vp, real, real, vp]
            getattr(L, "or_fast_linear_interp_knots" + suf).restype = real
            getattr(L, "or_fast_linear_interp_knots" + suf).argtypes = [real, vp, ctypes.c_uint32, real, real]
            getattr(L, "or_fast_linear_interp_sorted" + suf).restype = real
            getattr(L, "or_fast_linear_interp_sorted" + suf).argtypes = [real, vp, vp, ctypes.c_uint32]
            getattr(L, "or_emit" + suf).argtypes = [vp, vp, vp, ctypes.c_uint64, ctypes.c_uint32,
                                                    ctypes.c_uint64, vp]
            getattr(L, "or_propagate" + suf).restype = ctypes.c_int64
            getattr(L, "or_propagate" + suf).argtypes = [vp, vp, vp, vp, ctypes.c_uint64, ctypes.c_uint32,
                                                        ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64,
                                                        vp, ctypes.c_uint64, vp, ctypes.c_int]
        L.or_philox4x32_10.argtypes = [vp, vp, vp]
        L.or_uniform.restype = ctypes.c_float
        L.or_uniform.argtypes = [ctypes.c_uint32]
        L.or_block_draws.argtypes = [vp, vp, vp]
        L.or_rodrigues_rotation.argtypes = [vp, vp, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _real(precision):
    return (np.float32, "_f32") if precision == 32 else (np.float64, "_f64")


def philox4x32_10(counter, key):
    c = np.ascontiguousarray(counter, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def uniform(word):
    return lib().or_uniform(int(word))


def block_draws(words):
    """(u[4] on (0, 1], p[2] azimuths in revolutions) of one Philox block."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    u = np.zeros(4, dtype=np.float32)
    p = np.zeros(2, dtype=np.float32)
    lib().or_block_draws(_p(w), _p(u), _p(p))
    return u, p


def hg_costheta(g, eta, precision=32):
    _, suf = _real(precision)
    return getattr(lib(), "or_hg_costheta" + suf)(g, eta)


def rotate_to_axis(old_dir, new_dir, precision=32):
    dt, suf = _real(precision)
    a = np.ascontiguousarray(old_dir, dtype=dt)
    b = np.ascontiguousarray(new_dir, dtype=dt)
    out = np.zeros(3, dtype=dt)
    getattr(lib(), "or_rotate_to_axis" + suf)(_p(a), _p(b), _p(out))
    return out


def rodrigues_rotation(a, b, operand):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    v = np.ascontiguousarray(operand, dtype=np.float64)
    out = np.zeros(3)
    lib().or_rodrigues_rotation(_p(a), _p(b), _p(v), _p(out))
    return out


def update_position(pos, direction, step, precision=32):
    dt, suf = _real(precision)
    p = np.ascontiguousarray(pos, dtype=dt)
    d = np.ascontiguousarray(direction, dtype=dt)
    out = np.zeros(3, dtype=dt)
    getattr(lib(), "or_update_position" + suf)(_p(p), _p(d), step, _p(out))
    return out


def ray_sphere(pos, direction, center, radius, step, precision=32):
    dt, suf = _real(precision)
    p = np.ascontiguousarray(pos, dtype=dt)
    d = np.ascontiguousarray(direction, dtype=dt)
    c = np.ascontiguousarray(center, dtype=dt)
    out = np.zeros(1, dtype=dt)
    hit = getattr(lib(), "or_ray_sphere" + suf)(_p(p), _p(d), _p(c), dt(radius) * dt(radius), step, _p(out))
    return bool(hit), float(out[0])


def fast_linear_interp_knots(x, knots, lower, upper, precision=32):
    dt, suf = _real(precision)
    k = np.ascontiguousarray(knots, dtype=dt)
    return getattr(lib(), "or_fast_linear_interp_knots" + suf)(x, _p(k), k.shape[0], lower, upper)


def fast_linear_interp_sorted(x, xs, ys, precision=32):
    dt, suf = _real(precision)
    a = np.ascontiguousarray(xs, dtype=dt)
    b = np.ascontiguousarray(ys, dtype=dt)
    return getattr(lib(), "or_fast_linear_interp_sorted" + suf)(x, _p(a), _p(b), a.shape[0])


def emit(tables, spectrum_struct, emitter_struct, seed, stream, photon_ids, precision=32):
    """Emission draws of the given photon ids: array (n, 4) of {λ, dx, dy, dz}."""
    dt, suf = _real(precision)
    ids = np.asarray(photon_ids, dtype=np.uint64)
    out = np.zeros((ids.shape[0], 4), dtype=dt)
    fn = getattr(lib(), "or_emit" + suf)
    for i, pid in enumerate(ids):
        fn(ctypes.byref(tables), ctypes.byref(spectrum_struct), ctypes.byref(emitter_struct),
           int(seed), int(stream), int(pid), ctypes.c_void_p(out[i:].ctypes.data))
    return out


def propagate(tables, spectrum_struct, emitter_struct, sphere_struct, seed=0, stream=0,
              max_steps=10, id_begin=0, id_end=None, hits_cap=None, n_threads=None,
              precision=32, hit_dtype=None, summary_struct=None):
    """Run the oracle over photon ids [id_begin, id_end).  `tables` is a
    pp_medium_tables ctypes struct; the other structs are the ctypes structs of
    include/photonprop.h.  Returns (hits, summary struct)."""
    if id_end is None:
        id_end = int(emitter_struct.n_photons)
    n = max(0, id_end - id_begin)
    cap = int(n if hits_cap is None else hits_cap)
    hits = np.zeros(max(cap, 1), dtype=hit_dtype)
    summ = summary_struct()
    if n_threads is None:
        n_threads = len(os.sched_getaffinity(0))
    _, suf = _real(precision)
    getattr(lib(), "or_propagate" + suf)(
        ctypes.byref(tables), ctypes.byref(spectrum_struct), ctypes.byref(emitter_struct),
        ctypes.byref(sphere_struct), int(seed), int(stream), int(max_steps), int(id_begin),
        int(id_end), _p(hits), cap, ctypes.byref(summ), int(n_threads))
    return hits[: summ.n_stored], summ
