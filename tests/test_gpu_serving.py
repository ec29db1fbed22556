"""Serving loop: continuous batching into a persistent kernel.

Each request's result must be the same alone (Context.propagate) and in the
mix (submitted to the server while others are on the GPU): summaries bitwise,
hit sets bitwise by photon id.
"""
import time

import numpy as np
import pytest

import photonprop as pp
from photonprop import workloads

pytestmark = pytest.mark.gpu
SUM_KEYS = ("n_sim", "n_steps", "n_hits", "sum_w", "sum_wt", "sum_wt2")


def say(*a):
    print(*a, flush=True)


def requests(n, seed):
    reqs = workloads.batched_pairs(n, seed=seed, photons=(1e3, 2e6), distance=(0.8, 5.0))
    if n > 3:
        reqs["n_photons"][3] = 0
    return reqs


def alone(ctx, r):
    src = pp.PhotonSource(tuple(r["pos"]), tuple(r["axis"]), float(r["t0"]), int(r["n_photons"]),
                          pp.InverseSquareSpectrum(300.0, 600.0),
                          pp.AngularEmissionProfile("CherenkovCone"), float(r["beta"]))
    tgt = pp.DetectionSphere(tuple(r["center"]), float(r["radius"]))
    return ctx.propagate(src, tgt, seed=int(r["seed"]), stream=int(r["stream"]),
                         max_steps=int(r["max_steps"]))


def key(h):
    h = np.sort(h, order="photon_id").copy()
    h["request"] = 0
    return h


def test_server_results_equal_alone(ctx):
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    reqs = requests(40, seed=21)
    caps = reqs["n_photons"].astype(np.uint64)
    say("serving: 40 mixed requests, max_slots 8 (so later ones queue while earlier run)")
    with pp.Server(ctx, spec, max_slots=8, max_tiles_per_request=512,
                   hit_pool_records=1 << 24) as srv:
        tickets = []
        for i, r in enumerate(reqs):
            tickets.append(srv.submit(r, int(caps[i])))
            if i % 5 == 4:
                time.sleep(0.002)  # arrivals while others are on the GPU
        got = []
        for i, t in enumerate(tickets):
            assert srv.wait(t, 60000), "ticket %d timed out" % i
            got.append(srv.result(t, int(caps[i])))
        st = srv.stats()
    say("server stats", st)
    assert st["completed"] == st["submitted"] == len(reqs)
    assert st["kernel_launches"] >= 1
    for i, r in enumerate(reqs):
        s, h, lat = got[i]
        assert lat > 0
        if i in (0, 1, 2, 3, 7, 11, 19, 39):
            ref = alone(ctx, r)
            for k in SUM_KEYS:
                assert s[k] == ref.summary[k], (i, k, s[k], ref.summary[k])
            np.testing.assert_array_equal(key(h), key(ref.hits))
        assert s["n_sim"] == r["n_photons"]
        assert s["n_stored"] == h.shape[0] == min(s["n_hits"], caps[i])


def test_server_large_requests_in_chunks_equal_alone(ctx):
    """Requests above 512 tiles (4.2e6 photons) are published in chunks,
    first come first served, while small ones go out whole ahead of them; a
    request's chunks interleave with other requests' tiles.  Every result is
    still bitwise the one it has alone."""
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    reqs = workloads.batched_pairs(8, seed=44, photons=(1e3, 1e4), distance=(6.0, 12.0))
    reqs["n_photons"] = [6_000_000, 20_000, 7_500_000, 3_000, 5_000_000, 400_000, 9_000_000, 70_000]
    caps = [20_000] * len(reqs)
    say("serving: 4 large (chunked) and 4 small requests, max_slots 6")
    with pp.Server(ctx, spec, max_slots=6, max_tiles_per_request=2048,
                   hit_pool_records=1 << 20) as srv:
        tickets = [srv.submit(r, caps[i]) for i, r in enumerate(reqs)]
        got = []
        for i, t in enumerate(tickets):
            assert srv.wait(t, 60000), "ticket %d timed out" % i
            got.append(srv.result(t, caps[i]))
        st = srv.stats()
    say("server stats", st)
    assert st["tiles_published"] == sum((int(n) + 8191) // 8192 for n in reqs["n_photons"])
    for i, r in enumerate(reqs):
        s, h, _ = got[i]
        ref = alone(ctx, r)
        for k in SUM_KEYS:
            assert s[k] == ref.summary[k], (i, k, s[k], ref.summary[k])
        assert s["n_hits"] <= caps[i], "cap too small for an exact hit comparison"
        np.testing.assert_array_equal(key(h), key(ref.hits))


def test_server_overflow_and_summary_only(ctx):
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    r = requests(1, seed=2)[0]
    r["n_photons"] = 300_000
    r["center"] = (1.2, 0.0, 0.0)
    r["axis"] = (1.0, 0.0, 1.0)
    ref = alone(ctx, r)
    assert ref.summary["n_hits"] > 10
    with pp.Server(ctx, spec, max_slots=4, max_tiles_per_request=64, hit_pool_records=1 << 16) as srv:
        t1 = srv.submit(r, 5)
        t2 = srv.submit(r, 0)
        assert srv.wait(t1, 60000) and srv.wait(t2, 60000)
        s1, h1, _ = srv.result(t1, 5)
        s2, h2, _ = srv.result(t2, 0)
        with pytest.raises(pp.PhotonPropError):
            big = r.copy()
            big["n_photons"] = 64 * 8192 + 1
            srv.submit(big, 0)
        with pytest.raises(pp.PhotonPropError):
            srv.submit(r, 1 << 17)  # above the pool
    assert s1["status"] == pp._abi.PP_ERR_OVERFLOW and s1["n_stored"] == 5 and h1.shape[0] == 5
    assert s2["n_stored"] == 0 and h2.shape[0] == 0
    for k in SUM_KEYS:
        assert s1[k] == ref.summary[k] == s2[k]


def test_server_open_loop_poisson(ctx):
    """A short Poisson arrival run of the serving workload shape (budgets
    log-uniform 1e4..1e7 here to keep the test short); every request finishes
    and latencies are recorded."""
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    arr, reqs = workloads.serving_requests(400, 2000.0, seed=5, photons=(1e4, 1e7))
    with pp.Server(ctx, spec, max_slots=256, max_tiles_per_request=2048,
                   hit_pool_records=1 << 20) as srv:
        out = pp.run_open_loop(srv, arr, reqs, hits_cap=0)
        st = srv.stats()
    lat = out["latency_us"]
    say("open loop 2k req/s: p50 %.0f us p99 %.0f us, %.3g photons/s, launches %d"
        % (np.percentile(lat, 50), np.percentile(lat, 99), out["photons"] / out["makespan_s"],
           st["kernel_launches"]))
    assert np.all(lat > 0)
    assert out["steps"] > 9 * out["photons"]


def test_server_open_loop_native_driver(ctx):
    """The C++ load generator drives the same schedule through the same
    entry points: every request finishes, and the photon steps (a property of
    the requests alone) equal the Python driver's."""
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    arr, reqs = workloads.serving_requests(400, 5000.0, seed=6, photons=(1e3, 1e6))
    with pp.Server(ctx, spec, max_slots=256, max_tiles_per_request=2048,
                   hit_pool_records=1 << 20) as srv:
        a = pp.run_open_loop_native(srv, arr, reqs, hits_cap=0)
        b = pp.run_open_loop(srv, arr, reqs, hits_cap=0)
    say("native open loop 5k req/s: p50 %.0f us p99 %.0f us, makespan %.3f s"
        % (np.percentile(a["latency_us"], 50), np.percentile(a["latency_us"], 99), a["makespan_s"]))
    assert np.all(a["latency_us"] > 0)
    assert a["steps"] == b["steps"] > 9 * a["photons"]
    # the same schedule as one pp_propagate call per request, from native code:
    # every request's photon steps equal the server's
    c = pp.run_per_request_native(ctx, spec, arr, reqs)
    say("native per-request loop 5k req/s: p50 %.0f us p99 %.0f us, makespan %.3f s"
        % (np.percentile(c["latency_us"], 50), np.percentile(c["latency_us"], 99), c["makespan_s"]))
    assert np.all(c["latency_us"] > 0)
    assert np.array_equal(c["steps_per_request"], a["steps_per_request"])


def test_server_entry_ring_wraps(ctx):
    """40,000 one-tile requests publish more than twice the 16,384 range
    entries of the ring (host copy and device copy alike): every request's
    photon steps equal those of the same request in one batched launch, so
    each tile found its own entry across the wraps."""
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    arr, reqs = workloads.serving_requests(40_000, 200_000.0, seed=8, photons=(500, 2000))
    with pp.Server(ctx, spec, max_slots=1024, max_tiles_per_request=64,
                   hit_pool_records=1 << 16) as srv:
        out = pp.run_open_loop_native(srv, arr, reqs, hits_cap=0, timeout_s=60.0)
        st = srv.stats()
    _, _, sums = ctx.propagate_batch(reqs, spec, hits_cap=1, allow_overflow=True)
    say("ring wrap: %d requests in %.3f s, %d tiles published"
        % (len(reqs), out["makespan_s"], st["tiles_published"]))
    assert st["completed"] == len(reqs)
    np.testing.assert_array_equal(out["steps_per_request"], sums["n_steps"])


def test_server_queued_request_does_not_hold_the_kernel(ctx):
    """A request waiting for pool room does not keep the persistent kernel
    running: with A holding the whole pool and done, the kernel idles out
    (other kernels run meanwhile), collecting A's result admits B, and B runs
    on a relaunched kernel with the same result as alone."""
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    reqs = requests(2, seed=55)
    reqs["n_photons"] = [100_000, 120_000]
    pool = 1 << 16
    with pp.Server(ctx, spec, max_slots=4, max_tiles_per_request=64, hit_pool_records=pool,
                   idle_exit_us=2000) as srv:
        ta = srv.submit(reqs[0], pool)
        tb = srv.submit(reqs[1], pool)  # waits: A holds the whole pool
        assert srv.wait(ta, 60000)
        time.sleep(0.05)  # well past idle_exit_us
        st = srv.stats()
        assert st["pending"] == 1 and st["in_flight"] == 0, st
        other = alone(ctx, reqs[0])  # another kernel on the device meanwhile
        sa, ha, _ = srv.result(ta, pool)
        assert srv.wait(tb, 60000)
        sb, hb, _ = srv.result(tb, pool)
        st = srv.stats()
    say("queued-request stats", st)
    assert st["kernel_launches"] >= 2
    for k in SUM_KEYS:
        assert sa[k] == other.summary[k], (k, sa[k], other.summary[k])
    ref = alone(ctx, reqs[1])
    for k in SUM_KEYS:
        assert sb[k] == ref.summary[k], (k, sb[k], ref.summary[k])
    np.testing.assert_array_equal(key(hb), key(ref.hits))


def test_server_idle_stop_and_relaunch(ctx):
    """The kernel stops after idle_exit_us without work and is relaunched on
    the next submission; tickets resume at the published count, so results
    after the relaunch are still the same as alone."""
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    reqs = requests(6, seed=33)
    reqs["n_photons"] = [200_000, 50_000, 1_000_000, 8192, 8193, 300_000]
    with pp.Server(ctx, spec, max_slots=4, max_tiles_per_request=256, hit_pool_records=1 << 20,
                   idle_exit_us=2000) as srv:
        got = []
        for wave in (reqs[:3], reqs[3:]):
            ts = [srv.submit(r, int(r["n_photons"])) for r in wave]
            for t, r in zip(ts, wave):
                assert srv.wait(t, 60000)
                got.append(srv.result(t, int(r["n_photons"])))
            time.sleep(0.05)  # well past idle_exit_us: the kernel stops
        st = srv.stats()
    say("idle/relaunch stats", st)
    assert st["kernel_launches"] >= 2
    for r, (s, h, _) in zip(reqs, got):
        ref = alone(ctx, r)
        for k in SUM_KEYS:
            assert s[k] == ref.summary[k], (k, s[k], ref.summary[k])
        np.testing.assert_array_equal(key(h), key(ref.hits))


def test_server_zero_photon_request_takes_no_slot(ctx):
    """A request without photons completes at once and takes neither a slot
    nor pool records.  On a fresh server, collecting it and letting the worker
    poll leaves completed == submitted and nothing in flight (the slot it would
    have held is not reaped as a phantom completion); the ticket is then
    unknown, and two requests submitted next each equal Context.propagate
    alone."""
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    reqs = requests(3, seed=71)
    reqs["n_photons"] = [0, 150_000, 90_000]
    pool = 1 << 18
    with pp.Server(ctx, spec, max_slots=2, max_tiles_per_request=64, hit_pool_records=pool) as srv:
        t0 = srv.submit(reqs[0], 1000)
        assert srv.wait(t0, 10000)
        s0, h0, _ = srv.result(t0, 1000)
        assert s0["n_sim"] == 0 and s0["n_hits"] == 0 and h0.shape[0] == 0
        time.sleep(0.03)  # several worker passes
        st = srv.stats()
        say("after the zero-photon request:", st)
        assert st["submitted"] == st["completed"] == 1
        assert st["in_flight"] == 0 and st["pending"] == 0 and st["pool_records_free"] == pool
        with pytest.raises(pp.PhotonPropError) as e:
            srv.wait(t0, 10)
        assert e.value.code == pp._abi.PP_ERR_NOT_FOUND
        ta = srv.submit(reqs[1], 150_000)
        tb = srv.submit(reqs[2], 90_000)
        assert srv.wait(ta, 60000) and srv.wait(tb, 60000)
        got = [srv.result(ta, 150_000), srv.result(tb, 90_000)]
        st = srv.stats()
    assert st["submitted"] == st["completed"] == 3
    for r, (s, h, _) in zip(reqs[1:], got):
        ref = alone(ctx, r)
        for k in SUM_KEYS:
            assert s[k] == ref.summary[k], (k, s[k], ref.summary[k])
        np.testing.assert_array_equal(key(h), key(ref.hits))


@pytest.mark.parametrize("rate", [2000.0, 5000.0])
def test_server_serving_workload_full_size(ctx, rate):
    """The serving workload at its stated size: 400 requests, Poisson arrivals
    at `rate`/s, budgets log-uniform on 1e4..1e8 photons, random emitter pose,
    targets (r = 0.21 m) at log-uniform 1..150 m.  Every request completes
    (p50/p99 reported).  Nine requests keep their hits -- the four largest
    budgets beyond 20 m (all above 1e7 photons), four across the budget range
    and one of up to 3e6 photons whose target sits on the Cherenkov cone: their summaries
    are bitwise, and their hit sets by photon id equal, those of pp_propagate
    alone.  The last of them is also checked photon by photon against the
    oracle."""
    from oracle.compare import assert_parity, compare_hits, oracle_for

    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    arr, reqs = workloads.serving_requests(400, rate, seed=17, photons=(1e4, 1e8),
                                           distance=(1.0, 150.0))
    n = len(reqs)
    dist = np.linalg.norm(reqs["center"], axis=1)
    big = [i for i in np.argsort(reqs["n_photons"])[::-1] if dist[i] > 20.0][:4]
    order = np.argsort(reqs["n_photons"])
    spread = [int(order[j]) for j in (0, 120, 240, 360)]
    # the oracle's request: up to 3e6 photons, the target within 4° of the
    # Cherenkov cone (~42°) and the largest photons x solid angle among those
    cone = np.degrees(np.arccos(np.clip(np.einsum("ij,ij->i", reqs["axis"].astype(np.float64),
                                                  reqs["center"].astype(np.float64)) / dist, -1, 1)))
    small = np.nonzero(reqs["n_photons"] <= 3_000_000)[0]
    score = reqs["n_photons"][small] * (0.21 / dist[small]) ** 2 * (np.abs(cone[small] - 42.0) < 4.0)
    near = int(small[np.argmax(score)])
    pick = sorted(set(int(i) for i in big) | set(spread) | {near})
    assert len(pick) == 9 and min(int(reqs["n_photons"][i]) for i in big) > 10_000_000
    caps = np.zeros(n, dtype=np.uint64)
    caps[pick] = 1 << 20
    say("serving workload at %d/s: %d requests, %.3g photons; kept hits of %s (photons %s)"
        % (rate, n, float(reqs["n_photons"].sum()), pick, [int(reqs["n_photons"][i]) for i in pick]))
    with pp.Server(ctx, spec, max_slots=1024, max_tiles_per_request=16384,
                   hit_pool_records=1 << 24) as srv:
        out = pp.run_open_loop(srv, arr, reqs, hits_cap=caps, timeout_s=100.0)
        st = srv.stats()
    lat = out["latency_us"]
    say("serving workload %d/s: p50 %.2f ms p99 %.2f ms, %.3g photons/s, launches %d"
        % (rate, np.percentile(lat, 50) / 1e3, np.percentile(lat, 99) / 1e3,
           out["photons"] / out["makespan_s"], st["kernel_launches"]))
    assert st["submitted"] == st["completed"] == n
    assert all(s is not None and s["n_sim"] == r["n_photons"] for s, r in zip(out["summaries"], reqs))
    for i in pick:
        r = reqs[i]
        s, h = out["summaries"][i], out["hits"][i]
        src = pp.PhotonSource(tuple(r["pos"]), tuple(r["axis"]), float(r["t0"]), int(r["n_photons"]),
                              spec, pp.AngularEmissionProfile("CherenkovCone"), float(r["beta"]))
        tgt = pp.DetectionSphere(tuple(r["center"]), float(r["radius"]))
        ref = ctx.propagate(src, tgt, seed=int(r["seed"]), stream=int(r["stream"]),
                            max_steps=int(r["max_steps"]), hits_cap=1 << 20)
        assert s["n_hits"] <= (1 << 20), "cap too small for an exact hit comparison"
        for k in SUM_KEYS:
            assert s[k] == ref.summary[k], (i, k, s[k], ref.summary[k])
        np.testing.assert_array_equal(key(h), key(ref.hits))
        if i == near:
            orc, osum = oracle_for(ctx, src, tgt, seed=int(r["seed"]), stream=int(r["stream"]))
            cmp = compare_hits(h, orc)
            say("request %d against the oracle:" % i, cmp)
            assert osum["n_hits"] > 1000
            assert_parity(cmp, osum["n_hits"])


def test_server_kernel_stops_itself_after_device_wait(ctx):
    """The kernel's own stop: with the host's idle stop 10 s away, the kernel
    stops by itself once it has waited device_wait_us = 3 ms with nothing
    published.  It is then no longer running, the stop is counted, and a
    request submitted next runs on a relaunched kernel (resumed at the tile
    the kernel reported) with the same result as alone."""
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    reqs = requests(2, seed=88)
    reqs["n_photons"] = [250_000, 180_000]
    with pp.Server(ctx, spec, max_slots=4, max_tiles_per_request=64, hit_pool_records=1 << 20,
                   idle_exit_us=10_000_000, device_wait_us=3000) as srv:
        ta = srv.submit(reqs[0], 250_000)
        assert srv.wait(ta, 60000)
        sa, ha, _ = srv.result(ta, 250_000)
        time.sleep(0.1)  # far beyond 3 ms, far below the host's 10 s
        st = srv.stats()
        say("after the device wait:", st)
        assert st["device_stops"] >= 1 and st["kernel_running"] == 0
        tb = srv.submit(reqs[1], 180_000)
        assert srv.wait(tb, 60000)
        sb, hb, _ = srv.result(tb, 180_000)
        st = srv.stats()
    assert st["kernel_launches"] >= 2 and st["completed"] == 2
    for r, s, h in ((reqs[0], sa, ha), (reqs[1], sb, hb)):
        ref = alone(ctx, r)
        for k in SUM_KEYS:
            assert s[k] == ref.summary[k], (k, s[k], ref.summary[k])
        np.testing.assert_array_equal(key(h), key(ref.hits))


def test_server_host_pool_results_equal_alone(ctx):
    """The hit pool in pinned host memory (PP_SERVER_HOST_POOL; the mode used
    without an SDMA engine): the kernel writes the records straight to the
    host, no device-to-host copy is involved, and every result equals the same
    request alone."""
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    reqs = requests(12, seed=91)
    caps = [int(n) for n in reqs["n_photons"]]
    with pp.Server(ctx, spec, max_slots=4, max_tiles_per_request=512, hit_pool_records=1 << 22,
                   host_pool=True) as srv:
        assert srv.stats()["host_pool"] == 1
        tickets = [srv.submit(r, caps[i]) for i, r in enumerate(reqs)]
        got = []
        for i, t in enumerate(tickets):
            assert srv.wait(t, 60000), i
            got.append(srv.result(t, caps[i]))
        st = srv.stats()
    say("host-pool stats", st)
    assert st["completed"] == len(reqs)
    for i, r in enumerate(reqs):
        s, h, _ = got[i]
        ref = alone(ctx, r)
        for k in SUM_KEYS:
            assert s[k] == ref.summary[k], (i, k, s[k], ref.summary[k])
        np.testing.assert_array_equal(key(h), key(ref.hits))


def test_server_leave_room_grid_and_results(ctx):
    """PP_SERVER_LEAVE_ROOM: the persistent grid takes at most 7 of 8 wave
    slots per SIMD (a smaller grid than the default, which fills the device),
    and requests still equal the same request alone."""
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    reqs = requests(6, seed=93)
    caps = [int(n) for n in reqs["n_photons"]]
    grids = {}
    for leave in (False, True):
        with pp.Server(ctx, spec, max_slots=4, max_tiles_per_request=512,
                       hit_pool_records=1 << 22, leave_room=leave) as srv:
            tickets = [srv.submit(r, caps[i]) for i, r in enumerate(reqs)]
            got = []
            for i, t in enumerate(tickets):
                assert srv.wait(t, 60000), i
                got.append(srv.result(t, caps[i]))
            grids[leave] = srv.stats()["grid_blocks"]
        for i, r in enumerate(reqs):
            s, h, _ = got[i]
            ref = alone(ctx, r)
            for k in SUM_KEYS:
                assert s[k] == ref.summary[k], (leave, i, k)
            np.testing.assert_array_equal(key(h), key(ref.hits))
    _, n_cu = ctx.device_info()
    say("grid blocks: full %d, leave_room %d (%d CUs)" % (grids[False], grids[True], n_cu))
    assert grids[True] < grids[False]
    assert grids[True] <= 7 * n_cu


def test_server_concurrent_submitters_and_collectors(ctx):
    """Four host threads submit, wait for and collect their own requests on one
    server at once (the C calls release the interpreter lock, so they overlap),
    while a fifth polls the statistics: every request completes, each equals
    the same request alone, and the counters add up."""
    import threading

    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    per, n_thr = 12, 4
    reqs = workloads.batched_pairs(per * n_thr, seed=97, photons=(1e3, 3e6), distance=(0.8, 5.0))
    caps = [int(n) for n in reqs["n_photons"]]
    got = [None] * len(reqs)
    errors = []
    stop = threading.Event()
    # the pool holds every request's records at once: a collector waiting for
    # its last request must not wait for pool room that only its own
    # uncollected results hold
    pool = 1 << 26
    assert sum(caps) <= pool
    with pp.Server(ctx, spec, max_slots=8, max_tiles_per_request=512,
                   hit_pool_records=pool) as srv:
        def worker(w):
            try:
                mine = list(range(w, len(reqs), n_thr))
                tickets = [(i, srv.submit(reqs[i], caps[i])) for i in mine]
                for i, t in reversed(tickets):  # collect in the opposite order
                    assert srv.wait(t, 60000), i
                    got[i] = srv.result(t, caps[i])
            except Exception as e:  # reported on the main thread
                errors.append(e)

        def poller():
            try:
                while not stop.is_set():
                    st = srv.stats()
                    assert st["completed"] <= st["submitted"], st
                    time.sleep(0.001)
            except Exception as e:
                errors.append(e)

        threads = [threading.Thread(target=worker, args=(w,)) for w in range(n_thr)]
        p = threading.Thread(target=poller)
        p.start()
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        stop.set()
        p.join()
        st = srv.stats()
    assert not errors, errors
    say("concurrent submitters: stats", st)
    assert st["submitted"] == st["completed"] == len(reqs) and st["in_flight"] == 0
    assert st["pool_records_free"] == pool
    for i, r in enumerate(reqs):
        s, h, _ = got[i]
        ref = alone(ctx, r)
        for k in SUM_KEYS:
            assert s[k] == ref.summary[k], (i, k)
        np.testing.assert_array_equal(key(h), key(ref.hits))
