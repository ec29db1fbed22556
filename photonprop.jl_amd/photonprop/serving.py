"""Serving loop: continuous batching of propagation requests into a
persistent GPU kernel (pp_server_* in include/photonprop.h).

Requests arrive with uneven photon budgets; the library admits them as device
slots and hit-pool room (the HBM budget) free up, and publishes their photon
tiles to a kernel that keeps running while work arrives.  A request's result
is bitwise the same as Context.propagate gives for it alone.
"""
import ctypes
import threading
import time

import numpy as np

from . import _abi
from .propagation import REQUEST_DTYPE, spectrum_struct, summary_dict


class Server:
    def __init__(self, ctx, spectrum, max_slots=256, max_tiles_per_request=16384,
                 hit_pool_records=1 << 24, idle_exit_us=20000, device_wait_us=0,
                 host_pool=False, leave_room=False):
        """device_wait_us: the kernel stops by itself after waiting this long with
        nothing published (0: the library's default); host_pool: keep the hit
        pool in pinned host memory (PP_SERVER_HOST_POOL); leave_room: leave a
        wave slot and 32 VGPRs per SIMD for other kernels (PP_SERVER_LEAVE_ROOM)."""
        self.ctx = ctx
        self.lib = ctx.lib
        cfg = _abi.pp_server_config(max_slots, max_tiles_per_request, int(hit_pool_records),
                                    int(idle_exit_us), int(device_wait_us),
                                    (_abi.PP_SERVER_HOST_POOL if host_pool else 0) |
                                    (_abi.PP_SERVER_LEAVE_ROOM if leave_room else 0), 0)
        spec, keep = spectrum_struct(spectrum)
        self._h = ctypes.c_void_p()
        _abi.check(self.lib.pp_server_create(ctx.handle, ctypes.byref(spec), ctypes.byref(cfg),
                                             ctypes.byref(self._h)), self.lib)
        del keep

    def submit(self, request, hits_cap=0):
        """Queue one REQUEST_DTYPE record; returns a ticket."""
        r = np.ascontiguousarray(np.asarray(request, dtype=REQUEST_DTYPE).reshape(1))
        t = ctypes.c_uint64()
        _abi.check(self.lib.pp_server_submit(self._h, r.ctypes.data_as(ctypes.POINTER(_abi.pp_request)),
                                             int(hits_cap), ctypes.byref(t)), self.lib)
        return t.value

    def wait(self, ticket, timeout_ms=10000):
        done = ctypes.c_int()
        _abi.check(self.lib.pp_server_wait(self._h, int(ticket), int(timeout_ms), ctypes.byref(done)),
                   self.lib)
        return bool(done.value)

    def result(self, ticket, hits_cap=0):
        """(summary dict, hits, latency in µs) of a finished ticket; frees it."""
        hits = np.empty(max(int(hits_cap), 0), dtype=_abi.HIT_DTYPE)
        s = _abi.pp_summary()
        n = ctypes.c_uint64()
        lat = ctypes.c_double()
        _abi.check(self.lib.pp_server_result(self._h, int(ticket), ctypes.byref(s), _abi.ptr(hits),
                                             int(hits_cap), ctypes.byref(n), ctypes.byref(lat)),
                   self.lib)
        return summary_dict(s), hits[: n.value], lat.value

    def stats(self):
        st = _abi.pp_server_stats()
        _abi.check(self.lib.pp_server_stats_get(self._h, ctypes.byref(st)), self.lib)
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self._h:
            self.lib.pp_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_open_loop(server, arrivals, requests, hits_cap=0, timeout_s=120.0):
    """Submit requests[i] at arrivals[i] seconds (Poisson schedule) from one
    thread while another collects results in order.  hits_cap is one capacity
    for every request or one per request.  Returns a dict with per-request
    latencies in µs from the scheduled arrival to completion (the library's
    submit-to-completion time plus any lag of the submitting thread behind the
    schedule), the makespan, the photon / step totals, every request's summary
    and the hits of the requests with a nonzero capacity (by index)."""
    n = len(requests)
    caps = np.broadcast_to(np.asarray(hits_cap, dtype=np.uint64), (n,))
    tickets = [None] * n
    lag = np.zeros(n)
    lat = np.zeros(n)
    steps = np.zeros(n, dtype=np.uint64)
    summaries = [None] * n
    kept = {}
    submitted = threading.Semaphore(0)
    err = []

    def collect():
        try:
            for i in range(n):
                submitted.acquire()
                while not server.wait(tickets[i], 1000):
                    if time.perf_counter() - t0 > timeout_s:
                        raise TimeoutError("request %d not finished" % i)
                s, h, l = server.result(tickets[i], int(caps[i]))
                lat[i] = l
                steps[i] = s["n_steps"]
                summaries[i] = s
                if caps[i]:
                    kept[i] = h
        except Exception as e:  # surfaced after join
            err.append(e)

    th = threading.Thread(target=collect)
    t0 = time.perf_counter()
    th.start()
    for i in range(n):
        delay = arrivals[i] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        lag[i] = 1e6 * max(0.0, time.perf_counter() - t0 - arrivals[i])
        tickets[i] = server.submit(requests[i], int(caps[i]))
        submitted.release()
    th.join()
    makespan = time.perf_counter() - t0
    if err:
        raise err[0]
    return {
        "latency_us": lat + lag,
        "submit_lag_us": lag,
        "makespan_s": makespan,
        "photons": int(np.asarray(requests["n_photons"], dtype=np.uint64).sum()),
        "steps": int(steps.sum()),
        "summaries": summaries,
        "hits": kept,
    }


_LOADGEN = None
LOADGEN_TIMEOUT = -100
LOADGEN_BAD_ARG = -101


def _loadgen():
    """libpploadgen.so (built beside libphotonprop.so; raises if missing)."""
    global _LOADGEN
    if _LOADGEN is None:
        import os
        path = os.path.join(os.path.dirname(_abi.LIB_PATH), "libpploadgen.so")
        if not os.path.exists(path):
            raise OSError("libpploadgen.so not found at %s; run __graft_entry__.build()" % path)
        lg = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        lg.ppl_open_loop.restype = ctypes.c_int
        lg.ppl_open_loop.argtypes = [vp, vp, vp, vp, vp, dp, ctypes.c_uint64, ctypes.c_uint64,
                                     ctypes.c_double, dp, dp, ctypes.POINTER(ctypes.c_uint64), dp]
        lg.ppl_per_request_loop.restype = ctypes.c_int
        lg.ppl_per_request_loop.argtypes = [vp, vp, vp, vp, dp, ctypes.c_uint64, ctypes.c_double,
                                            dp, ctypes.POINTER(ctypes.c_uint64), dp]
        _LOADGEN = lg
    return _LOADGEN


def run_open_loop_native(server, arrivals, requests, hits_cap=0, timeout_s=120.0):
    """run_open_loop with the submitting and collecting threads in C++
    (libpploadgen.so, through the same pp_server_* entry points of the library
    instance that created `server`).  Same return dict."""
    lg = _loadgen()
    lib = server.lib
    reqs = np.ascontiguousarray(requests, dtype=REQUEST_DTYPE)
    arr = np.ascontiguousarray(arrivals, dtype=np.float64)
    n = reqs.shape[0]
    lat = np.zeros(n)
    lag = np.zeros(n)
    steps = np.zeros(n, dtype=np.uint64)
    mk = ctypes.c_double()
    fn = lambda f: ctypes.cast(f, ctypes.c_void_p)
    dptr = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = lg.ppl_open_loop(server._h, fn(lib.pp_server_submit), fn(lib.pp_server_wait),
                          fn(lib.pp_server_result), reqs.ctypes.data_as(ctypes.c_void_p), dptr(arr),
                          n, int(hits_cap), float(timeout_s), dptr(lat), dptr(lag),
                          steps.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(mk))
    if rc == LOADGEN_TIMEOUT:
        raise TimeoutError("open loop: requests not finished within %.0f s" % timeout_s)
    if rc == LOADGEN_BAD_ARG:
        raise ValueError("open loop: invalid arguments")
    _abi.check(rc, lib)
    return {
        "latency_us": lat,
        "submit_lag_us": lag,
        "makespan_s": mk.value,
        "photons": int(np.asarray(reqs["n_photons"], dtype=np.uint64).sum()),
        "steps": int(steps.sum()),
        "steps_per_request": steps,
    }


def run_per_request_native(ctx, spectrum, arrivals, requests, timeout_s=300.0):
    """The same open-loop schedule served by one pp_propagate call per request
    (first in, first out) from native code (libpploadgen.so's
    ppl_per_request_loop, through the entry point of the library instance that
    created `ctx`).  Returns {latency_us, makespan_s, photons, steps,
    steps_per_request}."""
    lg = _loadgen()
    lib = ctx.lib
    reqs = np.ascontiguousarray(requests, dtype=REQUEST_DTYPE)
    arr = np.ascontiguousarray(arrivals, dtype=np.float64)
    n = reqs.shape[0]
    lat = np.zeros(n)
    steps = np.zeros(n, dtype=np.uint64)
    mk = ctypes.c_double()
    spec, keep = spectrum_struct(spectrum)
    rc = lg.ppl_per_request_loop(ctx.handle, ctypes.cast(lib.pp_propagate, ctypes.c_void_p),
                                 ctypes.cast(ctypes.byref(spec), ctypes.c_void_p),
                                 reqs.ctypes.data_as(ctypes.c_void_p),
                                 arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
                                 float(timeout_s),
                                 lat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 steps.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                 ctypes.byref(mk))
    del keep
    if rc == LOADGEN_TIMEOUT:
        raise TimeoutError("per-request loop: not finished within %.0f s" % timeout_s)
    if rc == LOADGEN_BAD_ARG:
        raise ValueError("per-request loop: invalid arguments")
    _abi.check(rc, lib)
    return {
        "latency_us": lat,
        "makespan_s": mk.value,
        "photons": int(np.asarray(reqs["n_photons"], dtype=np.uint64).sum()),
        "steps": int(steps.sum()),
        "steps_per_request": steps,
    }
