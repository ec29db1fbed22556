"""The allocation-free batched entry point (pp_propagate_batch_device_noalloc)
against pp_propagate_batch_device, through the C ABI: two calls in a row, the
second with a smaller input, each with the same summaries (bitwise), offsets
and per-request hit records; sizes beyond the stated limits are refused
before anything runs."""
import ctypes

import numpy as np
import pytest

import photonprop as pp
from photonprop import _abi, workloads
from photonprop.propagation import REQUEST_DTYPE, spectrum_struct

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def run(ctx, name, reqs, spectrum, hits_cap, per_cap):
    """One call of entry point `name` with torch device buffers; returns
    (hits, offsets, summaries) on the host."""
    import torch

    reqs = np.ascontiguousarray(reqs, dtype=REQUEST_DTYPE)
    n = reqs.shape[0]
    out_cap = min(hits_cap, per_cap * n) if per_cap else hits_cap
    d_hits = torch.empty((max(out_cap, 1), 64), dtype=torch.uint8, device="cuda:0")
    d_off = torch.empty(n + 1, dtype=torch.int64, device="cuda:0")
    d_sum = torch.empty((n, 64), dtype=torch.uint8, device="cuda:0")
    torch.cuda.synchronize()
    spec, keep = spectrum_struct(spectrum)
    fn = getattr(ctx.lib, name)
    _abi.check(fn(ctx.handle, reqs.ctypes.data_as(ctypes.POINTER(_abi.pp_request)), n,
                  ctypes.byref(spec), ctypes.c_void_p(d_hits.data_ptr()), hits_cap, per_cap,
                  ctypes.c_void_p(d_off.data_ptr()), ctypes.c_void_p(d_sum.data_ptr())), ctx.lib)
    _abi.check(ctx.lib.pp_synchronize(ctx.handle), ctx.lib)
    del keep
    off = d_off.cpu().numpy().astype(np.uint64)
    sums = np.frombuffer(d_sum.cpu().numpy().tobytes(), dtype=_abi.SUMMARY_DTYPE)
    hits = np.frombuffer(d_hits.cpu().numpy().tobytes(), dtype=_abi.HIT_DTYPE)[: int(off[-1])]
    return hits, off, sums


def by_request(hits, off, r):
    h = hits[int(off[r]):int(off[r + 1])]
    return np.sort(h, order="photon_id")


def same(a, b, label):
    ha, oa, sa = a
    hb, ob, sb = b
    np.testing.assert_array_equal(oa, ob, err_msg=label)
    assert sa.tobytes() == sb.tobytes(), label
    for r in range(len(sa)):
        np.testing.assert_array_equal(by_request(ha, oa, r), by_request(hb, ob, r), err_msg=label)


def test_noalloc_batch_equals_allocating_batch(ctx):
    # call 1: 64 uneven requests (an empty one, one exactly one tile), 1/λ² spectrum
    r1 = workloads.batched_pairs(64, seed=5, photons=(1e3, 3e5), distance=(0.8, 6.0))
    r1["n_photons"][3] = 0
    r1["n_photons"][4] = 8192
    r1["max_steps"][::9] = 13
    s1 = pp.InverseSquareSpectrum(300.0, 600.0)
    cap1 = pp._abi.PP_NOALLOC_MAX_HITS
    say("no-alloc call 1: 64 requests, %d photons, hits_cap %d" % (int(r1["n_photons"].sum()), cap1))
    a = run(ctx, "pp_propagate_batch_device_noalloc", r1, s1, cap1, 0)
    b = run(ctx, "pp_propagate_batch_device", r1, s1, cap1, 0)
    say("call 1: %d hits stored, statuses %s" % (int(a[1][-1]), set(a[2]["status"].tolist())))
    assert int(a[1][-1]) == int(a[2]["n_hits"].sum()) > 1000 and np.all(a[2]["status"] == 0)
    np.testing.assert_array_equal(a[2]["n_sim"], r1["n_photons"])
    same(a, b, "call 1")
    # call 2, smaller: 9 requests, a monochromatic spectrum, a per-request cap
    r2 = workloads.batched_pairs(9, seed=6, photons=(1e3, 1e5), distance=(0.8, 3.0))
    s2 = pp.Monochromatic(420.0)
    say("no-alloc call 2: 9 requests, %d photons, hits_cap 5000, per-request cap 40"
        % int(r2["n_photons"].sum()))
    a2 = run(ctx, "pp_propagate_batch_device_noalloc", r2, s2, 5000, 40)
    b2 = run(ctx, "pp_propagate_batch_device", r2, s2, 5000, 40)
    assert len(a2[2]) == 9 and int(a2[1][-1]) == int(a2[2]["n_stored"].sum()) > 0
    np.testing.assert_array_equal(a2[2]["n_sim"], r2["n_photons"])
    np.testing.assert_array_equal(a2[2]["n_stored"], np.minimum(a2[2]["n_hits"], 40))
    # the records kept under the per-request cap are a subset that depends on
    # the order of the kernel's appends, so only the summaries, offsets and
    # counts are compared here, and the kept records must be hits of the same
    # request in the uncapped result
    np.testing.assert_array_equal(a2[1], b2[1])
    assert a2[2].tobytes() == b2[2].tobytes()
    full = run(ctx, "pp_propagate_batch_device_noalloc", r2, s2, 5000, 0)
    for r in range(9):
        kept = set(by_request(a2[0], a2[1], r)["photon_id"].tolist())
        assert kept <= set(by_request(full[0], full[1], r)["photon_id"].tolist())


def test_noalloc_refuses_sizes_beyond_its_limits(ctx):
    import torch

    spec, keep = spectrum_struct(pp.InverseSquareSpectrum(300.0, 600.0))
    d = torch.empty(1 << 12, dtype=torch.uint8, device="cuda:0")
    p = ctypes.c_void_p(d.data_ptr())

    def call(reqs, hits_cap):
        reqs = np.ascontiguousarray(reqs, dtype=REQUEST_DTYPE)
        return ctx.lib.pp_propagate_batch_device_noalloc(
            ctx.handle, reqs.ctypes.data_as(ctypes.POINTER(_abi.pp_request)), reqs.shape[0],
            ctypes.byref(spec), p, hits_cap, 0, p, p)

    many = workloads.batched_pairs(pp._abi.PP_NOALLOC_MAX_REQUESTS + 1, seed=1, photons=(10, 20))
    assert call(many, 16) == _abi.PP_ERR_INVALID_ARG
    assert b"PP_NOALLOC_MAX_REQUESTS" in ctx.lib.pp_last_error()
    one = workloads.batched_pairs(1, seed=1, photons=(10, 20))
    assert call(one, pp._abi.PP_NOALLOC_MAX_HITS + 1) == _abi.PP_ERR_INVALID_ARG
    assert b"PP_NOALLOC_MAX_HITS" in ctx.lib.pp_last_error()
    one["n_photons"] = pp._abi.PP_NOALLOC_MAX_TILES * pp._abi.PP_TILE_PHOTONS + 1
    assert call(one, 16) == _abi.PP_ERR_INVALID_ARG
    assert b"PP_NOALLOC_MAX_TILES" in ctx.lib.pp_last_error()
    del keep
