"""Benchmark: photon steps/s on one MI355X for the 64-geometry scale workload.

Workload (one "step"): a point emitter with a Cherenkov cone about +z
(λ ∝ 1/λ² on 300-600 nm, β = 1) in the seawater-like medium (n 1.36→1.34,
λ_abs 10/30/10 m at 300/450/600 nm, λ_sca = 50 m·(λ/450 nm)^4, g = 0.92),
1e9 photons towards each of 64 spheres (r = 0.21 m): 16 distances
geomspace(1, 100, 16) m × 4 directions at 0°, 45°, 90°, 135° from the emitter
axis, max_steps = 10, Philox streams keyed by (seed, geometry).  All 64
geometries run as one persistent batched launch; hits are compacted per
geometry in device memory.

Prints one JSON line (see the fields below).  --gpus N with N > 1 exits: this
build measures one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "photonprop.jl_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

FLOPS_PER_STEP = 81          # canonical FP32 flops per scattering-loop iteration
PEAK_FP32_TFLOPS = 157.3     # MI355X published FP32 peak
HIT_BYTES = 64
# HBM bytes per propagate_kernel launch of the default workload (64 geometries x
# 1e9 photons): rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE x 1, the factors from a
# 1 GiB device copy (profiles/propagate_kernel_64geom_1e9.md).  Reported only
# for that workload.
PROFILED_PHOTONS = 1e9
PROFILED_TRAFFIC_BYTES = 1.77844e10


def workload_requests(pp, photons_per_geometry, seed=0):
    """The 64 geometries: 16 distances geomspace(1, 100, 16) m x targets at
    0/45/90/135 degrees from the +z axis, r = 0.21 m, max_steps 10."""
    from photonprop import workloads

    return workloads.scale_geometries(photons=photons_per_geometry, radius=0.21, seed=seed)


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pp, photons_per_geometry_sample, runs=5):
    """The oracle (FP32, OpenMP over all host cores this process may use) on
    the same 64 geometries with a bounded photon sample per geometry: one
    warm-up run, then the median of `runs` timed runs."""
    import oracle
    from oracle.compare import oracle_for

    tables = pp.medium_tables(pp.SeawaterConfigMedium())
    reqs = workload_requests(pp, photons_per_geometry_sample)
    threads = len(os.sched_getaffinity(0))
    oracle.lib()

    def one_run():
        steps = hits = 0
        t0 = time.perf_counter()
        for r in reqs:
            src = pp.PhotonSource(tuple(r["pos"]), tuple(r["axis"]), 0.0, int(r["n_photons"]),
                                  pp.InverseSquareSpectrum(300.0, 600.0))
            tgt = pp.DetectionSphere(tuple(r["center"]), float(r["radius"]))
            _, s = oracle_for(tables, src, tgt, seed=int(r["seed"]), stream=int(r["stream"]),
                              n_threads=threads)
            steps += s["n_steps"]
            hits += s["n_hits"]
        return steps, hits, time.perf_counter() - t0

    one_run()  # warm-up
    results = [one_run() for _ in range(runs)]
    steps, hits = results[0][0], results[0][1]
    assert all(r[0] == steps and r[1] == hits for r in results)  # same work every run
    times = sorted(r[2] for r in results)
    dt = times[len(times) // 2]
    return {
        "value": steps / dt,
        "unit": "photon steps/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": cpu_model(),
        "host_cpus": os.cpu_count(),
        "sample": "64 geometries x %d photons (%d steps, %d hits): median of %d runs after one "
                  "warm-up, %.2f s (runs %s s); detected photons/s %.4g"
                  % (photons_per_geometry_sample, steps, hits, runs, dt,
                     ", ".join("%.2f" % t for t in times), hits / dt),
        "detected_photons_per_s": hits / dt,
        "steps_per_s_per_core": steps / dt / threads,
    }


def bench_batched(pp, ctx, spectrum, reps=3):
    """4096 source-target pairs of uneven size in one persistent launch."""
    from photonprop import workloads

    reqs = workloads.batched_pairs(4096, seed=0)
    ctx.propagate_batch(reqs, spectrum)  # warm-up
    t0 = time.perf_counter()
    ks = []
    for _ in range(reps):
        hits, off, sums = ctx.propagate_batch(reqs, spectrum)
        ks.append(ctx.last_kernel_ms())
    dt = (time.perf_counter() - t0) / reps
    steps = int(sums["n_steps"].sum())
    return {
        "pairs": 4096,
        "photons": int(reqs["n_photons"].sum()),
        "photon_steps": steps,
        "hits": int(sums["n_hits"].sum()),
        "kernel_ms": sum(ks) / len(ks),
        "call_ms_incl_d2h": 1e3 * dt,
        "photon_steps_per_s_kernel": steps / (1e-3 * sum(ks) / len(ks)),
        "photon_steps_per_s_call": steps / dt,
    }


def summarize(res, n_req):
    lat = res["latency_us"]
    return {
        "p50_ms": float(np.percentile(lat, 50)) / 1e3,
        "p99_ms": float(np.percentile(lat, 99)) / 1e3,
        "requests_per_s": n_req / res["makespan_s"],
        "photons_per_s": res["photons"] / res["makespan_s"],
        "photon_steps_per_s": res["steps"] / res["makespan_s"],
    }


MIXED_RATES = (500.0, 1000.0, 1500.0, 2000.0, 2400.0, 2800.0, 3000.0, 3200.0, 3400.0, 3600.0,
               3800.0, 5000.0, 5500.0, 6000.0, 6500.0, 7000.0, 20000.0)
SMALL_RATES = (2000.0, 5000.0, 7000.0, 8000.0, 9000.0, 10000.0, 20000.0, 40000.0, 60000.0,
               80000.0, 120000.0, 160000.0, 240000.0, 400000.0, 500000.0, 800000.0)


def _serving_sweep(pp, ctx, spectrum, n_req, rates, photons, label):
    """Open-loop Poisson sweep of one request mix through (a) the server,
    driven by the native load generator (libpploadgen.so: submitting and
    collecting threads in C++), and (b) one pp_propagate launch per request
    (FIFO, driven by the same library's native per-request loop).  Latency is
    measured from the scheduled arrival in both.  A rate is sustained when p99
    < 20 ms and at least 95 % of the offered rate is delivered."""
    from photonprop import workloads

    out = {"photons_per_request": list(photons), "requests_per_rate": n_req, "server": {},
           "per_request_launch": {}}
    for rate in rates:
        arr, reqs = workloads.serving_requests(n_req, rate, seed=11, photons=photons)
        out["per_request_launch"][str(int(rate))] = summarize(
            pp.run_per_request_native(ctx, spectrum, arr, reqs), n_req)
        print("%s per-request launch %d/s: %s" % (label, rate,
                                                   out["per_request_launch"][str(int(rate))]),
              file=sys.stderr, flush=True)
    with pp.Server(ctx, spectrum, max_slots=1024, max_tiles_per_request=16384,
                   hit_pool_records=1 << 20) as srv:
        for rate in rates:
            arr, reqs = workloads.serving_requests(n_req, rate, seed=11, photons=photons)
            out["server"][str(int(rate))] = summarize(
                pp.run_open_loop_native(srv, arr, reqs, hits_cap=0), n_req)
            print("%s server %d/s: %s" % (label, rate, out["server"][str(int(rate))]),
                  file=sys.stderr, flush=True)
        out["server_stats"] = srv.stats()

    def sustained(d):
        ok = [v["requests_per_s"] for r, v in d.items()
              if v["p99_ms"] < 20.0 and v["requests_per_s"] >= 0.95 * float(r)]
        return max(ok) if ok else None

    a, b = sustained(out["server"]), sustained(out["per_request_launch"])
    out["server_sustained_requests_per_s_p99_under_20ms"] = a
    out["per_request_sustained_requests_per_s_p99_under_20ms"] = b
    out["server_over_per_request_sustained"] = (a / b) if (a and b) else None
    out["server_saturated_requests_per_s"] = max(v["requests_per_s"] for v in out["server"].values())
    out["per_request_saturated_requests_per_s"] = max(
        v["requests_per_s"] for v in out["per_request_launch"].values())
    return out


def bench_serving(pp, ctx, spectrum, n_req):
    """The serving workload (budgets log-uniform 1e4..1e8 photons, random
    pose, target 1-150 m log-uniform, Poisson arrivals) and a small-request
    mix (1e3..1e5 photons) where per-launch overhead rather than photon work
    bounds the one-launch-per-request path."""
    return {
        "mixed_1e4_1e8": _serving_sweep(pp, ctx, spectrum, n_req, MIXED_RATES, (1e4, 1e8), "mixed"),
        "small_1e3_1e5": _serving_sweep(pp, ctx, spectrum, 4 * n_req, SMALL_RATES, (1e3, 1e5),
                                        "small"),
    }


def bench_surrogate(pp, ctx, n_rows, photons_per_row, epochs, members=4, max_log_std=0.1,
                    n_eval=1024, eval_target_hits=1000.0, eval_photons=(100_000_000, 2_000_000_000),
                    activation=None):
    """Train the BF16-MFMA surrogate (an ensemble of `members` networks on bootstrap
    resamples of the MC rows from the batched kernel), then evaluate it on
    held-out pairs against high-statistics MC: yield errors over every pair with
    a truth hit, next to the truth's own error bar, and over the pairs resolved
    to <= 5 %; the variance gate's fallback fraction and the error of the pairs
    it keeps; the hybrid's answers and their cost; the median Kolmogorov
    distance between the predicted arrival-time mixture and the MC residual
    times.  Speed: ensemble queries/s (forward of 1e6 rows per member, HIP
    events) against MC queries/s at the 4096-pair batch's mean budget.

    Photons per training row (and per held-out pair) come from a pilot
    network's yield estimate ŷ: clip(20 / ŷ, photons_per_row, 10·photons_per_row)
    for training, clip(eval_target_hits / ŷ, *eval_photons) for the truth."""
    import torch
    from scipy.special import gammainc

    from photonprop import surrogate as sg
    from photonprop import workloads

    out = {}
    t0 = time.perf_counter()
    pilot = sg.pilot_allocator(ctx, n_min=photons_per_row, n_max=10 * photons_per_row, seed=11,
                               activation=activation)
    ts = sg.generate_training_set(ctx, n_rows, n_times=16, seed=3, allocate=pilot)
    out["training_rows"] = n_rows
    out["training_photons"] = ts.photons
    out["photons_per_row"] = [photons_per_row, 10 * photons_per_row]
    out["datagen_s"] = time.perf_counter() - t0
    print("surrogate: %d rows (%.3g photons) generated in %.1f s" % (n_rows, ts.photons,
                                                                     out["datagen_s"]),
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ens, hists = sg.train_ensemble(ctx, ts, k=members, epochs=epochs, batch=4096, lr=3e-3,
                                   time_weight=0.5, seed=0, activation=activation)
    out["train_s"] = time.perf_counter() - t0
    out["members"] = members
    out["activation"] = ens.members[0].activation
    out["epochs"] = epochs
    out["final_loss"] = [h[-1] for h in hists]
    print("surrogate: %d networks trained in %.1f s" % (members, out["train_s"]), file=sys.stderr,
          flush=True)
    del ts
    truth_alloc = sg.YieldAllocator(pilot.model, target_hits=eval_target_hits,
                                    n_min=eval_photons[0], n_max=eval_photons[1])
    pairs = sg.held_out_pairs(ctx, n_eval, None, seed=1234, per_request_cap=2000,
                              allocate=truth_alloc)
    pilot.model.close()
    hy = sg.HybridYield(ens, ctx, max_log_std=max_log_std)
    _, rep = sg.evaluate_hybrid(hy, pairs)
    out["eval_photons_per_pair"] = {
        "rule": "clip(%g / pilot yield, %g, %g)" % (eval_target_hits, *eval_photons),
        "total": int(sum(int(p["reqs"]["n_photons"].sum()) for p in pairs))}
    out["eval"] = rep
    # the hybrid against plain MC at the same per-pair budget rule for every pair
    rate = rep["hybrid_mc_photons"] / rep["hybrid_mc_kernel_s"] if rep["hybrid_mc_kernel_s"] else None
    out["hybrid_queries_per_s"] = rep["pairs"] / rep["hybrid_s"]
    out["mc_same_rule_queries_per_s"] = (rep["pairs"] / (rep["all_mc_photons"] / rate)
                                         if rate else None)
    out["hybrid_speedup_vs_mc_same_rule"] = (out["hybrid_queries_per_s"] /
                                             out["mc_same_rule_queries_per_s"]
                                             if rate else None)
    ks_d = []
    for p in pairs:
        _, _, ys = ens.predict(sg.scale_features(p["raw"]))
        w, a, b = ens.time_mixture(ys)
        tg = sg.t_geo(np.linalg.norm(p["reqs"]["center"], axis=1), p["reqs"]["radius"], p["medium"])
        off = p["offsets"]
        for r in np.nonzero(p["sums"]["sum_w"] >= 400.0)[0]:
            t = np.sort(p["hits"]["time"][int(off[r]):int(off[r + 1])].astype(np.float64) - tg[r])
            t = np.maximum(t, 0.0) + 1e-2
            if t.size < 50:
                continue
            cdf = (w[r][None, :] * gammainc(a[r][None, :], b[r][None, :] * t[:, None])).sum(1)
            emp = np.arange(1, t.size + 1) / t.size
            ks_d.append(float(np.max(np.abs(cdf - emp))))
    out["median_time_ks_distance"] = float(np.median(ks_d)) if ks_d else None
    del pairs
    ctx.set_medium(pp.SeawaterConfigMedium())
    # speed: ensemble forward on 1e6 queries vs MC per query
    n_q = 1_000_000
    xq = torch.rand((n_q, 8), device="cuda:0") * 2 - 1
    yq = torch.empty((n_q, 10), device="cuda:0")
    torch.cuda.synchronize()  # the library's stream does not wait for torch's
    ms = 0.0
    for m in ens.members:
        m.forward_device(xq.data_ptr(), yq.data_ptr(), n_q)
        ms += min(m.forward_device(xq.data_ptr(), yq.data_ptr(), n_q) for _ in range(3))
    out["surrogate_queries_per_s"] = n_q / (1e-3 * ms)
    out["surrogate_tflops"] = (members * n_q * 2.0 * (8 * 256 + 3 * 256 * 256 + 256 * 10)
                               / (1e-3 * ms) / 1e12)
    reqs = workloads.batched_pairs(4096, seed=0)
    ctx.propagate_batch(reqs, pp.InverseSquareSpectrum(300.0, 600.0), hits_cap=0, allow_overflow=True)
    kms = ctx.last_kernel_ms()
    out["mc_queries_per_s"] = 4096 / (1e-3 * kms)
    out["mc_mean_photons_per_query"] = float(reqs["n_photons"].mean())
    out["speedup_vs_mc"] = out["surrogate_queries_per_s"] / out["mc_queries_per_s"]
    ens.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons", type=float, default=1e9, help="photons per geometry")
    ap.add_argument("--cpu-sample", type=float, default=2e6,
                    help="photons per geometry for the CPU baseline, timed as the median of 5 "
                         "runs after a warm-up (0 skips it)")
    ap.add_argument("--extras", type=int, default=1,
                    help="also measure the 4096-pair batch and the serving loop (0 skips)")
    ap.add_argument("--serve-requests", type=int, default=1000,
                    help="requests per arrival rate in the serving measurement")
    ap.add_argument("--surrogate-rows", type=float, default=2e6,
                    help="MC training rows for the surrogate measurement (0 skips it)")
    ap.add_argument("--surrogate-photons", type=float, default=1e5, help="photons per training row")
    ap.add_argument("--surrogate-epochs", type=int, default=40)
    ap.add_argument("--surrogate-members", type=int, default=4,
                    help="networks in the ensemble whose spread is the predictive uncertainty")
    ap.add_argument("--surrogate-activation", choices=("relu", "silu"), default=None,
                    help="hidden activation of the surrogate networks (default: the library's "
                         "Python default)")
    ap.add_argument("--mlp-probe", type=float, default=0,
                    help="only time the surrogate's fused forward on this many random rows "
                         "(random weights; for profiling) and exit")
    ap.add_argument("--copy-probe", type=float, default=0,
                    help="only run a device-to-device copy of this many bytes (HBM counter "
                         "calibration for rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    ap.add_argument("--serve-probe", type=float, default=0,
                    help="only run the 64 geometries at this many photons each through one "
                         "batched launch and through the serving loop (all submitted at once), "
                         "twice each, and exit (for profiling serve_kernel)")
    args = ap.parse_args()
    if args.serve_probe:
        import photonprop as pp
        from photonprop import workloads

        spec = pp.InverseSquareSpectrum(300.0, 600.0)
        ctx = pp.Context(pp.SeawaterConfigMedium(), device=0)
        reqs = workloads.scale_geometries(photons=int(args.serve_probe), seed=1)
        batch_ms, serve_ms = [], []
        for _ in range(2):
            _, _, sums = ctx.propagate_batch(reqs, spec, hits_cap=1, allow_overflow=True)
            batch_ms.append(ctx.last_kernel_ms())
        with pp.Server(ctx, spec, max_slots=256, max_tiles_per_request=16384,
                       hit_pool_records=1 << 20) as srv:
            for _ in range(2):
                t0 = time.perf_counter()
                ts = [srv.submit(r, 0) for r in reqs]
                for t in ts:
                    assert srv.wait(t, 60000)
                serve_ms.append(1e3 * (time.perf_counter() - t0))
                got = [srv.result(t, 0)[0] for t in ts]
                assert all(g["n_steps"] == sums["n_steps"][i] for i, g in enumerate(got))
            grid = srv.stats()["grid_blocks"]
        print(json.dumps({"serve_probe_photons": int(args.serve_probe), "batch_kernel_ms": batch_ms,
                          "server_wall_ms": serve_ms, "server_grid_blocks": grid}))
        ctx.close()
        return
    if args.copy_probe:
        import torch

        n = int(args.copy_probe) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 4 * n, "dispatches": 3}))
        return
    if args.mlp_probe:
        import torch

        import photonprop as pp
        from photonprop import surrogate as sg

        n = int(args.mlp_probe)
        ctx = pp.Context(pp.SeawaterConfigMedium(), device=0)
        mlp = sg.MLP(ctx, seed=0, activation=args.surrogate_activation)
        xq = torch.rand((n, 8), device="cuda:0") * 2 - 1
        yq = torch.empty((n, 10), device="cuda:0")
        torch.cuda.synchronize()  # the library's stream does not wait for torch's
        times = [mlp.forward_device(xq.data_ptr(), yq.data_ptr(), n) for _ in range(5)]
        fl = n * 2.0 * (8 * 256 + 3 * 256 * 256 + 256 * 10)
        print(json.dumps({"mlp_probe_rows": n, "activation": mlp.activation, "forward_ms": times,
                          "queries_per_s": n / (1e-3 * min(times)),
                          "tflops": fl / (1e-3 * min(times)) / 1e12}))
        mlp.close()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU; --gpus %d is not supported (the path would shard by "
              "photon-id range, one range per GPU)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    import photonprop as pp

    n_per = int(args.photons)
    reqs = workload_requests(pp, n_per)
    spectrum = pp.InverseSquareSpectrum(300.0, 600.0)
    ctx = pp.Context(pp.SeawaterConfigMedium(), device=0)
    # ~4.1e-3 hits per photon averaged over the 64 geometries; keep 2x headroom
    cap = int(max(1 << 20, 8.2e-3 * n_per * len(reqs)))
    d_hits = torch.empty((cap, HIT_BYTES), dtype=torch.uint8, device="cuda:0")
    d_off = torch.zeros(len(reqs) + 1, dtype=torch.int64, device="cuda:0")
    d_sum = torch.zeros((len(reqs), 64), dtype=torch.uint8, device="cuda:0")
    spec, keep = pp.propagation.spectrum_struct(spectrum)
    lib = ctx.lib
    creqs = reqs.ctypes.data_as(pp._abi.ctypes.POINTER(pp._abi.pp_request))

    def step():
        pp._abi.check(lib.pp_propagate_batch_device(
            ctx.handle, creqs, len(reqs), pp._abi.ctypes.byref(spec),
            pp._abi.ctypes.c_void_p(d_hits.data_ptr()), cap, 0,
            pp._abi.ctypes.c_void_p(d_off.data_ptr()), pp._abi.ctypes.c_void_p(d_sum.data_ptr())), lib)

    for i in range(args.warmup):
        step()
        pp._abi.check(lib.pp_synchronize(ctx.handle), lib)
        print("warmup %d done" % i, file=sys.stderr, flush=True)

    kernel_ms = []
    torch.cuda.synchronize()
    pp._abi.check(lib.pp_synchronize(ctx.handle), lib)
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        kernel_ms.append(ctx.last_kernel_ms())  # waits for this step's kernel
        print("step %d: kernel %.1f ms" % (i, kernel_ms[-1]), file=sys.stderr, flush=True)
    pp._abi.check(lib.pp_synchronize(ctx.handle), lib)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    sums = np.frombuffer(d_sum.cpu().numpy().tobytes(), dtype=pp._abi.SUMMARY_DTYPE)
    steps_per_launch = int(sums["n_steps"].sum())
    hits_per_launch = int(sums["n_hits"].sum())
    stored = int(sums["n_stored"].sum())
    assert stored == hits_per_launch, "hit buffer overflow in the benchmark"
    assert int(sums["n_sim"].sum()) == n_per * len(reqs)
    ms_per_step = 1e3 * elapsed / args.steps
    steps_per_s = steps_per_launch * args.steps / elapsed
    avg_kernel_s = 1e-3 * sum(kernel_ms) / len(kernel_ms)
    achieved = steps_per_launch * FLOPS_PER_STEP / avg_kernel_s / 1e12
    name, n_cu = ctx.device_info()

    out = {
        "metric": "photon steps/s + detected photons/s at 1 GPU; p50/p99 request latency; surrogate err",
        "value": steps_per_s,
        "unit": "photon steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "64 geometries (16 distances geomspace(1,100,16) m x 0/45/90/135 deg), "
                        "%.3g photons each, seawater-like medium, Cherenkov cone 300-600 nm, "
                        "sphere r=0.21 m, max_steps 10, one batched persistent launch" % n_per,
            "photons_per_step": n_per * len(reqs),
            "parallelism": "single GPU",
        },
        "detected_photons_per_s": hits_per_launch * args.steps / elapsed,
        "photons_per_s": n_per * len(reqs) * args.steps / elapsed,
        "photon_steps_per_launch": steps_per_launch,
        "hits_per_launch": hits_per_launch,
        "kernel_ms_avg": 1e3 * avg_kernel_s,
        "device": "%s, %d CUs" % (name, n_cu),
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": PROFILED_TRAFFIC_BYTES if n_per == int(PROFILED_PHOTONS) else None,
            "traffic_over_algorithmic": (PROFILED_TRAFFIC_BYTES / (hits_per_launch * HIT_BYTES)
                                         if n_per == int(PROFILED_PHOTONS) else None),
            "note": "FP32 vector ALU (not MFMA): %d flops per photon step x steps per launch / "
                    "average propagate_kernel duration (HIP events on the library's stream); "
                    "the kernel is bound by VALU issue, HBM carries only the 64-B hit records"
                    % FLOPS_PER_STEP,
        },
    }
    if args.extras:
        out["batched_4096"] = bench_batched(pp, ctx, spectrum)
        print("batched done", file=sys.stderr, flush=True)
        out["serving"] = bench_serving(pp, ctx, spectrum, args.serve_requests)
        print("serving done", file=sys.stderr, flush=True)
        if args.surrogate_rows > 0:
            out["surrogate"] = bench_surrogate(pp, ctx, int(args.surrogate_rows),
                                               int(args.surrogate_photons), args.surrogate_epochs,
                                               members=args.surrogate_members,
                                               activation=args.surrogate_activation)
            print("surrogate done", file=sys.stderr, flush=True)
    if args.cpu_sample > 0:
        cb = cpu_baseline(pp, int(args.cpu_sample))
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = steps_per_s / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
