"""Batched propagation (one persistent launch, many uneven requests).

The pin is "same alone and in the mix": a request's summary and hits are
bitwise identical whether it runs alone (pp_propagate) or inside any batch.
"""
import numpy as np
import pytest

import photonprop as pp
from oracle.compare import assert_parity, compare_hits, oracle_for
from photonprop import workloads

pytestmark = pytest.mark.gpu
SUM_KEYS = ("n_sim", "n_steps", "n_hits", "sum_w", "sum_wt", "sum_wt2")


def mixed_batch(n=48, seed=3):
    rng = np.random.default_rng(seed)
    reqs = workloads.batched_pairs(n, seed=seed, photons=(1e2, 4e5), distance=(0.8, 6.0))
    reqs["n_photons"][5] = 0          # empty request in the middle
    reqs["n_photons"][6] = 8192       # exactly one tile
    reqs["n_photons"][7] = 8193       # one tile + 1
    reqs["max_steps"][::7] = 17
    reqs["seed"] = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
    return reqs


def alone(ctx, r, spectrum):
    src = pp.PhotonSource(tuple(r["pos"]), tuple(r["axis"]), float(r["t0"]), int(r["n_photons"]),
                          spectrum, pp.AngularEmissionProfile("CherenkovCone"), float(r["beta"]))
    tgt = pp.DetectionSphere(tuple(r["center"]), float(r["radius"]))
    return ctx.propagate(src, tgt, seed=int(r["seed"]), stream=int(r["stream"]),
                         max_steps=int(r["max_steps"])), src, tgt


def strip(h):
    h = np.sort(h, order="photon_id").copy()
    h["request"] = 0
    return h


def test_batch_equals_alone_bitwise(ctx):
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    reqs = mixed_batch()
    hits, off, sums = ctx.propagate_batch(reqs, spec)
    assert off[0] == 0 and np.all(np.diff(off.astype(np.int64)) >= 0)
    assert int(off[-1]) == hits.shape[0] == int(sums["n_hits"].sum())
    for r in range(len(reqs)):
        seg = hits[int(off[r]):int(off[r + 1])]
        assert np.all(seg["request"] == r)
        assert sums["n_sim"][r] == reqs["n_photons"][r]
        assert sums["n_stored"][r] == sums["n_hits"][r] == seg.shape[0]
    assert sums["n_hits"][5] == 0 and sums["n_steps"][5] == 0
    for r in (0, 5, 6, 7, 13, 21, 47):
        res, _, _ = alone(ctx, reqs[r], spec)
        for k in SUM_KEYS:
            assert res.summary[k] == sums[k][r], (r, k, res.summary[k], sums[k][r])
        np.testing.assert_array_equal(strip(res.hits), strip(hits[int(off[r]):int(off[r + 1])]))


def test_batch_order_does_not_matter(ctx):
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    reqs = mixed_batch(32, seed=9)
    h1, o1, s1 = ctx.propagate_batch(reqs, spec)
    perm = np.random.default_rng(1).permutation(len(reqs))
    h2, o2, s2 = ctx.propagate_batch(reqs[perm], spec)
    for j, r in enumerate(perm):
        for k in SUM_KEYS:
            assert s2[k][j] == s1[k][r]
        np.testing.assert_array_equal(strip(h2[int(o2[j]):int(o2[j + 1])]),
                                      strip(h1[int(o1[r]):int(o1[r + 1])]))


def test_batch_requests_match_oracle(ctx):
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    reqs = mixed_batch(16, seed=4)
    hits, off, sums = ctx.propagate_batch(reqs, spec)
    for r in (1, 2, 3):
        _, src, tgt = alone(ctx, reqs[r], spec)
        ref, rs = oracle_for(ctx, src, tgt, seed=int(reqs["seed"][r]), stream=int(reqs["stream"][r]),
                             max_steps=int(reqs["max_steps"][r]))
        cmp = compare_hits(hits[int(off[r]):int(off[r + 1])], ref)
        assert_parity(cmp, rs["n_hits"])


def test_batch_overflow_is_exact(ctx):
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    reqs = mixed_batch(24, seed=5)
    _, _, full = ctx.propagate_batch(reqs, spec)
    total = int(full["n_hits"].sum())
    assert total > 100
    with pytest.raises(pp.PhotonPropError):
        ctx.propagate_batch(reqs, spec, hits_cap=total // 3)
    hits, off, part = ctx.propagate_batch(reqs, spec, hits_cap=total // 3, allow_overflow=True)
    assert hits.shape[0] == total // 3 == int(off[-1])
    for k in SUM_KEYS:
        np.testing.assert_array_equal(part[k], full[k])
    assert np.all(part["n_stored"] <= part["n_hits"])
    assert np.any(part["status"] == pp._abi.PP_ERR_OVERFLOW)


def test_batched_4096_pairs_workload(ctx):
    """The uneven 4096-pair workload in one launch: exact photon accounting and
    grouped output."""
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    reqs = workloads.batched_pairs(4096, seed=0)
    hits, off, sums = ctx.propagate_batch(reqs, spec)
    np.testing.assert_array_equal(sums["n_sim"], reqs["n_photons"])
    assert int(off[-1]) == int(sums["n_hits"].sum()) == hits.shape[0]
    assert np.all(sums["status"] == 0)
    steps = int(sums["n_steps"].sum())
    assert steps > 9.0 * int(reqs["n_photons"].sum())
    print("4096 pairs: photons %.3g steps %.3g hits %d, kernel %.1f ms"
          % (reqs["n_photons"].sum(), steps, hits.shape[0], ctx.last_kernel_ms()))


def test_per_request_cap_keeps_a_subset(ctx):
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    reqs = mixed_batch(20, seed=6)
    full_h, full_o, full_s = ctx.propagate_batch(reqs, spec)
    h, o, s = ctx.propagate_batch(reqs, spec, per_request_cap=3)
    for k in SUM_KEYS:
        np.testing.assert_array_equal(s[k], full_s[k])
    np.testing.assert_array_equal(s["n_stored"], np.minimum(full_s["n_hits"], 3))
    assert np.all(s["status"] == 0)
    for r in range(len(reqs)):
        seg = h[int(o[r]):int(o[r + 1])]
        full = full_h[int(full_o[r]):int(full_o[r + 1])]
        assert seg.shape[0] == min(3, full.shape[0])
        assert set(seg["photon_id"].tolist()) <= set(full["photon_id"].tolist())


def test_cascade_light_sources_batched_equal_alone():
    """The light-yield generator's requests (a 1 TeV electron cascade as point
    sources every 0.5 m along its axis, src/lightyield.jl:215-285) in the
    reference's cascadia water with its 20-knot Cherenkov spectrum, through
    one batched launch: every source's summary is bitwise, and its hit set by
    photon id equal to, the same source propagated alone, and each hit arrives
    no earlier than its source's delayed start t0."""
    from photonprop import lightyield as ly

    m = pp.make_cascadia_medium_properties()
    spec = pp.CherenkovSpectrum((300.0, 800.0), 20, m)
    p = ly.Particle((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 5.0, 1e3, ly.ParticleType.EMinus)
    tgt = pp.DetectionSphere((3.0, 0.0, 6.0), 0.3)
    reqs = ly.particle_to_elongated_lightsource(p, (0.0, 20.0), 0.5, m, (300.0, 800.0), tgt, seed=4)
    with pp.Context(m) as c:
        hits, off, sums = c.propagate_batch(reqs, spec)
        print("cascade: %d sources, %.3g photons, %d hits" % (len(reqs), float(reqs["n_photons"].sum()),
                                                             hits.shape[0]), flush=True)
        np.testing.assert_array_equal(sums["n_sim"], reqs["n_photons"])
        assert np.all(sums["status"] == 0) and int(sums["n_hits"].sum()) > 1000
        for r in range(len(reqs)):
            seg = hits[int(off[r]):int(off[r + 1])]
            assert np.all(seg["time"] >= reqs["t0"][r])
        for r in range(0, len(reqs), 5):
            res, _, _ = alone(c, reqs[r], spec)
            for k in SUM_KEYS:
                assert res.summary[k] == sums[k][r], (r, k, res.summary[k], sums[k][r])
            np.testing.assert_array_equal(strip(res.hits), strip(hits[int(off[r]):int(off[r + 1])]))
