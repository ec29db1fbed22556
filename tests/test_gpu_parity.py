"""GPU parity: the HIP library against the CPU oracle, through the C ABI.

Per-photon: both sides draw photon i from the same Philox stream, so the hit
sets must agree photon by photon (see oracle/compare.py for the tolerances).
Statistical: independent seeds, KS / χ² on time, wavelength and impact angle,
summed weight within 3σ.  Deterministic unit checks on the device helpers.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import stats

import oracle
import photonprop as pp
from oracle.compare import assert_parity, compare_hits, oracle_for

from closed_forms import C_LIGHT, beta_below, check_unscattered_hits

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)


def cone_source(n, spectrum=None, axis=(0.0, 0.0, 1.0), profile="CherenkovCone", beta=1.0):
    return pp.PhotonSource((0.0, 0.0, 0.0), axis, 0.0, int(n),
                           spectrum or pp.InverseSquareSpectrum(300.0, 600.0),
                           pp.AngularEmissionProfile(profile), beta)


# ------------------------------------------------------------ device helpers ----

def test_device_philox_matches_known_answers(ctx):
    kat = json.load(open(os.path.join(HERE, "golden", "philox_kat.json")))
    for case in kat["cases"]:
        ctr = np.array(case["counter"], dtype=np.uint32)
        key = case["key"][0] | (case["key"][1] << 32)
        out = np.zeros(4, dtype=np.uint32)
        pp._abi.check(ctx.lib.pp_selftest_philox(ctx.handle, pp._abi.ptr(ctr), key, pp._abi.ptr(out), 1))
        assert list(out) == case["out"], case
    # bulk: device == oracle on 4096 counters
    rng = np.random.default_rng(0)
    ctr = rng.integers(0, 2 ** 32, size=(4096, 4), dtype=np.uint64).astype(np.uint32)
    out = np.zeros((4096, 4), dtype=np.uint32)
    seed = 0x0123456789ABCDEF
    pp._abi.check(ctx.lib.pp_selftest_philox(ctx.handle, pp._abi.ptr(ctr), seed, pp._abi.ptr(out), 4096))
    for i in range(0, 4096, 97):
        want = oracle.philox4x32_10(ctr[i], [seed & 0xFFFFFFFF, seed >> 32])
        assert list(out[i]) == list(want)


def test_device_rotation_kats_and_rodrigues(ctx):
    ey = [0.0, 1.0, 0.0]
    old = np.array([[0, 0, 1], [0, 0, -1], ey, [0.2, 0.3, 1.0], [0, 0, -1]], dtype=np.float32)
    old[3] /= np.linalg.norm(old[3])
    new = np.array([ey, ey, ey, [-5.0, 1.0, -20.0], [0.48, -0.6, 0.64]], dtype=np.float32)
    new[3] /= np.linalg.norm(new[3])
    out = np.zeros_like(old)
    pp._abi.check(ctx.lib.pp_selftest_rotate(ctx.handle, pp._abi.ptr(old), pp._abi.ptr(new),
                                             pp._abi.ptr(out), 5))
    # src/tests/test_photon_prob.jl:26-30, :37-41, :43-53
    np.testing.assert_allclose(out[0], ey, atol=1e-7)
    np.testing.assert_allclose(out[1], [0, -1, 0], atol=1e-7)
    np.testing.assert_allclose(out[2], [0, 0, -1], atol=1e-6)
    assert abs(np.linalg.norm(out[3]) - 1) < 1e-6
    # the exact south pole with every component of the local direction nonzero:
    # the reference's (x, -y, -z) (src/photon_prop_cuda.jl:185-189)
    np.testing.assert_allclose(out[4], [0.48, 0.6, -0.64], atol=2e-7)
    # random pairs against Rodrigues (:55-64) and the FP64 oracle
    rng = np.random.default_rng(4)
    n = 20000
    a = rng.normal(size=(n, 3)); a /= np.linalg.norm(a, axis=1, keepdims=True)
    v = rng.normal(size=(n, 3)); v /= np.linalg.norm(v, axis=1, keepdims=True)
    a32, v32 = a.astype(np.float32), v.astype(np.float32)
    out = np.zeros_like(a32)
    pp._abi.check(ctx.lib.pp_selftest_rotate(ctx.handle, pp._abi.ptr(a32), pp._abi.ptr(v32),
                                             pp._abi.ptr(out), n))
    ez = np.array([0.0, 0.0, 1.0])
    worst = 0.0
    for i in range(0, n, 37):
        ai = a32[i].astype(np.float64); ai /= np.linalg.norm(ai)
        rr = oracle.rodrigues_rotation(ez, ai, v32[i].astype(np.float64))
        worst = max(worst, float(np.abs(out[i] - rr).max()))
    assert worst < 5e-6, worst
    # the angle to the old direction is the local polar angle, everywhere
    cosang = np.einsum("ij,ij->i", out.astype(np.float64), a32.astype(np.float64))
    assert np.abs(cosang - v32[:, 2]).max() < 5e-6
    # old directions within 1e-7 .. 1e-1 rad of either pole: the rotation's
    # 1/(1+z) is taken as 2/((1+z)² + x² + y²), which must stay accurate there
    m = 4000
    t = 10.0 ** rng.uniform(-7, -1, m)
    ph = rng.uniform(0.0, 2.0 * np.pi, m)
    sgn = np.where(rng.uniform(size=m) < 0.5, -1.0, 1.0)
    a = np.stack([t * np.cos(ph), t * np.sin(ph), sgn * np.sqrt(1.0 - t * t)], -1)
    v = rng.normal(size=(m, 3)); v /= np.linalg.norm(v, axis=1, keepdims=True)
    a32, v32 = a.astype(np.float32), v.astype(np.float32)
    out = np.zeros_like(a32)
    pp._abi.check(ctx.lib.pp_selftest_rotate(ctx.handle, pp._abi.ptr(a32), pp._abi.ptr(v32),
                                             pp._abi.ptr(out), m))
    worst = 0.0
    for i in range(0, m, 7):
        ai = a32[i].astype(np.float64); ai /= np.linalg.norm(ai)
        rr = oracle.rodrigues_rotation(ez, ai, v32[i].astype(np.float64))
        worst = max(worst, float(np.abs(out[i] - rr).max()))
    assert worst < 5e-6, worst


@pytest.mark.parametrize("g", [0.92, 0.97, 0.5])
def test_device_hg_mean_equals_g(ctx, g):
    # North-star unit check: mean HG cosine equals g to 1e-3 (1e8 device draws),
    # and E[cos²θ] = (1 + 2g²)/3.
    n = 100_000_000
    sums = np.zeros(2)
    pp._abi.check(ctx.lib.pp_selftest_hg_moments(ctx.handle, float(g), 11, n, pp._abi.ptr(sums)))
    mean, m2 = sums[0] / n, sums[1] / n
    say("HG g=%g: mean %.6f  E[c^2] %.6f (want %.6f)" % (g, mean, m2, (1 + 2 * g * g) / 3))
    assert abs(mean - g) < 1e-3
    assert abs(m2 - (1 + 2 * g * g) / 3) < 1e-3


def test_device_hg_per_draw_matches_oracle_and_cdf(ctx):
    g = 0.92
    n = 50_000
    out = np.zeros(n, dtype=np.float32)
    pp._abi.check(ctx.lib.pp_selftest_hg(ctx.handle, g, 3, pp._abi.ptr(out), n))
    words = np.concatenate([oracle.philox4x32_10([i, 0, 0, 0], [3, 0]) for i in range(n // 4)])
    # the propagation's draw: eta = 1 - (w >> 9)·2^-23 (exact in FP32)
    ref = np.array([oracle.hg_costheta(g, 1.0 - (int(w) >> 9) * 2.0 ** -23) for w in words])
    # same algebra, FP32; the device's reciprocal is the 1-ulp v_rcp_f32
    np.testing.assert_allclose(out, ref, atol=3e-6)
    cdf = lambda mu: (1 - g * g) / (2 * g) * (1 / np.sqrt(1 + g * g - 2 * g * mu) - 1 / (1 + g))
    assert stats.kstest(out.astype(np.float64), cdf).pvalue > 0.01


# --------------------------------------------------------- per-photon parity ----

CASES = [
    # (name, source kwargs, target centre, radius, max_steps, seed, stream)
    ("cone_near", dict(n=300_000), (1.5, 0.0, 1.5), 0.3, 10, 1, 0),
    ("cone_10m_90deg", dict(n=2_000_000), (10.0, 0.0, 0.0), 0.21, 10, 0, 0),
    ("cone_axis_tilted", dict(n=400_000, axis=(0.5, -0.3, 0.81)), (2.0, 0.5, 2.0), 0.5, 10, 5, 3),
    # emitter axis at the exact south pole: the emission rotation's undetermined case
    ("cone_axis_down", dict(n=300_000, axis=(0.0, 0.0, -1.0)), (1.5, 0.3, -1.5), 0.3, 10, 3, 0),
    ("mono_many_steps", dict(n=200_000, spectrum=pp.Monochromatic(420.0)), (0.0, 3.0, 0.5), 0.5, 37, 9, 1),
    ("uniform_sphere", dict(n=300_000, profile="UniformSphere"), (-2.0, 1.0, 0.0), 0.5, 10, 2, 0),
    ("ref_isotropic", dict(n=300_000, profile="IsotropicEmission"), (0.0, 0.0, 3.0), 0.4, 10, 0, 0),
    ("one_step", dict(n=300_000), (1.0, 0.0, 0.9), 0.4, 1, 4, 0),
    # the step loop's edges: the lone final step from block 0's spare word
    # (cone, even), one pair ending on the final step (cone, odd), and a final
    # pair / lone step that need their own Philox block (not the cone)
    ("two_steps_cone", dict(n=300_000), (1.0, 0.2, 0.9), 0.4, 2, 6, 0),
    ("three_steps_cone", dict(n=300_000), (0.9, -0.3, 1.0), 0.4, 3, 7, 2),
    ("five_steps_isotropic", dict(n=300_000, profile="IsotropicEmission"), (0.0, 1.0, 1.5), 0.4, 5,
     10, 0),
    ("four_steps_isotropic", dict(n=300_000, profile="IsotropicEmission"), (1.0, 1.0, 0.0), 0.4, 4,
     11, 1),
    ("ragged_tile", dict(n=8192 * 3 + 77), (1.2, 0.0, 1.1), 0.5, 10, 8, 0),
    ("tiny", dict(n=5), (0.5, 0.0, 0.45), 0.3, 10, 0, 0),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_per_photon_parity(ctx, case):
    name, skw, centre, radius, max_steps, seed, stream = case
    src = cone_source(**skw)
    tgt = pp.DetectionSphere(centre, radius)
    say("case", name, skw.get("n"), centre, radius, max_steps, seed, stream)
    res = ctx.propagate(src, tgt, seed=seed, stream=stream, max_steps=max_steps)
    ref, rsum = oracle_for(ctx, src, tgt, seed=seed, stream=stream, max_steps=max_steps)
    cmp = compare_hits(res.hits, ref)
    say(name, res.summary, cmp)
    s = res.summary
    assert s["n_sim"] == src.photons
    assert s["n_hits"] == s["n_stored"] == res.hits.shape[0]
    assert_parity(cmp, rsum["n_hits"])
    flips = cmp["only_gpu"] + cmp["only_cpu"]
    assert abs(int(s["n_steps"]) - int(rsum["n_steps"])) <= max_steps * flips
    if rsum["n_hits"]:
        # the FP64 weight and time moments, each to the same flip-based tolerance
        tol = 1e-4 + 2.0 * flips / rsum["n_hits"]
        for k in ("sum_w", "sum_wt", "sum_wt2"):
            assert s[k] == pytest.approx(rsum[k], rel=tol), (k, s[k], rsum[k])
    # every hit lies on the sphere surface and has a unit direction
    if res.hits.shape[0]:
        r = np.linalg.norm(res.hits["pos"].astype(np.float64) - np.array(centre), axis=1)
        assert np.abs(r - radius).max() < 1e-3
        assert np.abs(np.linalg.norm(res.hits["dir"], axis=1) - 1).max() < 1e-5
        assert np.all(res.hits["weight"] > 0) and np.all(res.hits["weight"] <= 1)


def test_photon_ids_across_two_to_the_32(ctx):
    """A request of 2^32 + 5e6 photons: the Philox counter's high id word and
    the 64-bit tile bases come into play.  The hits of photon ids in
    [2^32 - 5e6, 2^32 + 5e6) agree photon by photon with the oracle run on that
    id range alone, and the counts are exact."""
    n = (1 << 32) + 5_000_000
    a = math.radians(45.0)
    tgt = pp.DetectionSphere((21.54 * math.sin(a), 0.0, 21.54 * math.cos(a)), 0.21)
    src = cone_source(n)
    res = ctx.propagate(src, tgt, seed=3, stream=1, hits_cap=2_000_000)
    s = res.summary
    say("2^32+5e6 photons:", s)
    assert s["n_sim"] == n and s["n_stored"] == s["n_hits"] > 100_000
    assert 9.0 * n < s["n_steps"] <= 10 * n
    lo, hi = (1 << 32) - 5_000_000, (1 << 32) + 5_000_000
    ids = res.hits["photon_id"]
    mine = res.hits[(ids >= lo) & (ids < hi)]
    ref, rsum = oracle_for(ctx, src, tgt, seed=3, stream=1, id_begin=lo, id_end=hi)
    cmp = compare_hits(mine, ref)
    say("window [2^32 - 5e6, 2^32 + 5e6):", cmp)
    assert rsum["n_hits"] > 300 and np.any(ref["photon_id"] >= (1 << 32))
    assert_parity(cmp, rsum["n_hits"])


def test_knot_spectrum_and_reference_medium_parity():
    # The reference's own setup: cascadia water, CherenkovSpectrum((300, 800), 20),
    # :IsotropicEmission, g = 0.99, sphere on the z axis (src/photon_prop_cuda.jl:588-609).
    m = pp.make_cascadia_medium_properties()
    with pp.Context(m) as c2:
        spec = pp.CherenkovSpectrum((300.0, 800.0), 20, m)
        src = cone_source(500_000, spectrum=spec, profile="IsotropicEmission")
        tgt = pp.DetectionSphere((0.0, 0.0, 3.0), 0.21)
        res = c2.propagate(src, tgt, seed=0)
        ref, rsum = oracle_for(c2, src, tgt, seed=0)
        cmp = compare_hits(res.hits, ref)
        say("reference medium", res.summary, cmp)
        assert rsum["n_hits"] > 100
        assert_parity(cmp, rsum["n_hits"])
        df = pp.propagate_distance(3.0, m, 200_000)
        assert list(df.columns) == ["tres", "initial_theta", "ref_ix", "abs_weight",
                                    "dist_travelled", "wavelength"]
        assert len(df) > 10
        assert df["initial_theta"].max() <= 1.0 + 1e-6  # θ ~ U(-1, 1) rad in the reference profile
        assert (df["tres"] > -1.0).all()


def test_seed_and_stream_change_the_photons(ctx):
    src = cone_source(200_000)
    tgt = pp.DetectionSphere((1.5, 0.0, 1.5), 0.3)
    a = ctx.propagate(src, tgt, seed=1).hits
    b = ctx.propagate(src, tgt, seed=2).hits
    c = ctx.propagate(src, tgt, seed=1, stream=1).hits
    a2 = ctx.propagate(src, tgt, seed=1).hits
    np.testing.assert_array_equal(np.sort(a["photon_id"]), np.sort(a2["photon_id"]))
    ids = lambda h: set(h["photon_id"].tolist())
    assert len(ids(a) & ids(b)) < 0.2 * len(ids(a))
    assert len(ids(a) & ids(c)) < 0.2 * len(ids(a))


# ------------------------------------------------------ closed forms on device ----

class ConstMedium(pp.medium.MediumProperties):
    def __init__(self, n=1.34, abs_len=20.0, sca_len=1e30, g=0.92):
        self.n, self.a, self.s, self.hg_g = n, abs_len, sca_len, g
        self.table_range = (300.0, 600.0)
        self.table_points = 3

    def refractive_index(self, w):
        return np.full_like(np.asarray(w, dtype=float), self.n)

    def absorption_length(self, w):
        return np.full_like(np.asarray(w, dtype=float), self.a)

    def scattering_length(self, w):
        return np.full_like(np.asarray(w, dtype=float), self.s)


def test_isotropic_scattering_parity():
    # g = 0: isotropic scattering, a separate instance of the step loop;
    # per photon against the oracle in a medium that scatters every ~2 m.
    with pp.Context(ConstMedium(n=1.34, abs_len=20.0, sca_len=2.0, g=0.0)) as c:
        src = cone_source(300_000, spectrum=pp.Monochromatic(450.0))
        tgt = pp.DetectionSphere((1.5, 0.0, 1.5), 0.4)
        res = c.propagate(src, tgt, seed=6, stream=2)
        ref, rsum = oracle_for(c, src, tgt, seed=6, stream=2)
    cmp = compare_hits(res.hits, ref)
    say("isotropic scattering", res.summary, cmp)
    assert rsum["n_hits"] > 1000
    assert_parity(cmp, rsum["n_hits"])
    flips = cmp["only_gpu"] + cmp["only_cpu"]
    assert abs(int(res.summary["n_steps"]) - int(rsum["n_steps"])) <= 10 * flips


def test_device_survival_exp_minus_d_over_abs():
    # North-star unit check: survival equals exp(-d/λ_abs) to 1e-3.
    n, d, r, la = 1.34, 12.0, 0.21, 25.0
    with pp.Context(ConstMedium(n=n, abs_len=la)) as c:
        src = cone_source(1_000_000, spectrum=pp.Monochromatic(450.0), beta=1.0 / n)
        res = c.propagate(src, pp.DetectionSphere((0.0, 0.0, d), r))
    s = res.summary
    assert s["n_hits"] == src.photons
    surv = s["sum_w"] / src.photons
    say("survival", surv, math.exp(-(d - r) / la))
    assert abs(surv - math.exp(-(d - r) / la)) < 1e-3
    np.testing.assert_allclose(res.hits["dist"], d - r, rtol=1e-6)


def test_device_straight_line_hit_time_and_time_moments():
    """Photons emitted straight along the axis into a sphere on the axis: every
    hit's time is t0 + (d - r)·n/c (phase index, src/photon_prop_cuda.jl:634-638)
    and the FP64 moments are t·Σw and t²·Σw."""
    n, d, r, la, t0 = 1.34, 12.0, 0.21, 25.0, 17.25
    with pp.Context(ConstMedium(n=n, abs_len=la)) as c:
        src = pp.PhotonSource((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), t0, 1_000_000,
                              pp.Monochromatic(450.0), pp.AngularEmissionProfile("CherenkovCone"),
                              beta_below(n))
        res = c.propagate(src, pp.DetectionSphere((0.0, 0.0, d), r))
    s = res.summary
    L = d - r
    t = t0 + L * n / C_LIGHT
    say("straight line: time", float(res.hits["time"][0]), "want", t, s)
    assert s["n_hits"] == src.photons == res.hits.shape[0]
    np.testing.assert_allclose(res.hits["dist"], L, rtol=1e-6)
    np.testing.assert_allclose(res.hits["time"], t, rtol=1e-6)
    np.testing.assert_allclose(res.hits["weight"], math.exp(-L / la), rtol=1e-6)
    assert s["sum_w"] == pytest.approx(src.photons * math.exp(-L / la), rel=1e-6)
    assert s["sum_wt"] == pytest.approx(s["sum_w"] * t, rel=1e-6)
    assert s["sum_wt2"] == pytest.approx(s["sum_w"] * t * t, rel=1e-6)


def test_device_unscattered_hit_time_weight_and_moments(ctx):
    """Uniform emission in the configured seawater, one step allowed: every
    hit is an unscattered photon, whose time t0 + L·n(λ)/c, weight
    exp(-L/λ_abs(λ)) and path length follow from the geometry and the medium's
    formulas; the FP64 moments are the sums over the records."""
    centre, r, t0 = (1.1, -0.4, 2.3), 0.5, 3.75
    src = pp.PhotonSource((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), t0, 4_000_000,
                          pp.InverseSquareSpectrum(300.0, 600.0),
                          pp.AngularEmissionProfile("UniformSphere"))
    res = ctx.propagate(src, pp.DetectionSphere(centre, r), seed=12, max_steps=1)
    say("unscattered closed form:", res.summary)
    assert res.summary["n_hits"] > 20_000
    check_unscattered_hits(res.hits, res.summary, (0.0, 0.0, 0.0), centre, r, t0)


class RisingIndexMedium(ConstMedium):
    """n rising linearly from 1.33 at 300 nm to 1.35 at 800 nm (n_min at 300 nm)."""

    def __init__(self):
        super().__init__(n=1.33, abs_len=30.0)
        self.table_range = (300.0, 800.0)
        self.table_points = 501

    def refractive_index(self, w):
        return 1.33 + 0.02 * (np.asarray(w, dtype=float) - 300.0) / 500.0


def test_result_table_tres_straight_line():
    """result_table's tres (src/photon_prop_cuda.jl:639-641) for photons flying
    straight to a sphere on the axis at 300 nm: L = d - r, so
    tres = (d - r)·n(300)/c - (d - r)·n(800)/c, with the geometric time taken at
    n(800 nm); initial θ = 0 and ref_ix = n(300)."""
    m = RisingIndexMedium()
    d, r = 20.0, 0.21
    with pp.Context(m) as c:
        src = pp.PhotonSource((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0.0, 20_000, pp.Monochromatic(300.0),
                              pp.AngularEmissionProfile("CherenkovCone"), beta_below(1.33))
        res = c.propagate(src, pp.DetectionSphere((0.0, 0.0, d), r))
    assert res.summary["n_hits"] == src.photons
    df = pp.propagation.result_table(res.hits, m, d, r)
    want = (d - r) * 1.33 / C_LIGHT - (d - r) * 1.35 / C_LIGHT
    say("tres", float(df["tres"].iloc[0]), "want", want)
    # FP32 path length: 1e-7 of the ~88 ns photon time, i.e. ~1e-5 ns
    np.testing.assert_allclose(df["tres"], want, rtol=0, atol=5e-5)
    np.testing.assert_allclose(df["ref_ix"], 1.33, rtol=1e-12)
    np.testing.assert_allclose(df["initial_theta"], 0.0, atol=1e-3)
    np.testing.assert_allclose(df["dist_travelled"], d - r, rtol=1e-6)


def test_device_step_count_at_large_max_steps():
    """A target no photon can reach: every photon runs all max_steps = 2^20
    iterations, so n_steps = N·max_steps exactly (the per-tile step sums are
    64-bit above the lane level); max_steps above PP_MAX_STEPS is refused."""
    N, max_steps = 300, 1 << 20
    with pp.Context(ConstMedium(n=1.34, abs_len=20.0, sca_len=1.0)) as c:
        src = cone_source(N, spectrum=pp.Monochromatic(450.0))
        res = c.propagate(src, pp.DetectionSphere((1.0e6, 0.0, 0.0), 0.21), max_steps=max_steps)
        say("large max_steps:", res.summary)
        assert res.summary["n_hits"] == 0
        assert res.summary["n_steps"] == N * max_steps
        with pytest.raises(pp.PhotonPropError, match="max_steps"):
            c.propagate(src, pp.DetectionSphere((1.0, 0.0, 0.0), 0.21),
                        max_steps=pp._abi.PP_MAX_STEPS + 1)


def test_device_no_scatter_cone_fraction():
    n, d, r = 1.34, 5.0, 0.4
    thc = math.acos(1 / n)
    alpha = thc + 0.03
    with pp.Context(ConstMedium(n=n)) as c:
        src = cone_source(4_000_000, spectrum=pp.Monochromatic(450.0))
        res = c.propagate(src, pp.DetectionSphere((d * math.sin(alpha), 0, d * math.cos(alpha)), r))
    rho = math.asin(r / d)
    x = (math.cos(rho) - math.cos(thc) * math.cos(alpha)) / (math.sin(thc) * math.sin(alpha))
    p = math.acos(x) / math.pi
    N = src.photons
    k = res.summary["n_hits"]
    assert abs(k - N * p) < 5 * math.sqrt(N * p * (1 - p)), (k, N * p)


def test_device_unscattered_fraction_and_unit_weight():
    n, d, r, ls = 1.34, 6.0, 0.3, 8.0
    with pp.Context(ConstMedium(n=n, abs_len=1e30, sca_len=ls)) as c:
        src = cone_source(4_000_000, spectrum=pp.Monochromatic(450.0), beta=1.0 / n)
        res = c.propagate(src, pp.DetectionSphere((0.0, 0.0, d), r), max_steps=1)
    p = math.exp(-(d - r) / ls)
    N = src.photons
    k = res.summary["n_hits"]
    assert abs(k - N * p) < 5 * math.sqrt(N * p * (1 - p))
    assert np.all(res.hits["weight"] == 1.0)
    assert res.summary["n_steps"] == N


# ---------------------------------------------------------------- statistics ----

GEOMS = [(3.0, 45.0), (10.0, 45.0), (30.0, 45.0), (1.0, 0.0), (5.0, 90.0)]
OBSERVABLES = ("time", "wavelength", "impact")


def test_statistical_equivalence_independent_seeds(ctx):
    """Config physics at 5 target positions (distance m, angle from the emitter
    axis): GPU (2e8 photons, seed 1) against the FP64 oracle (seed 2, sized for
    >= ~3000 hits, up to 6e7 photons).  Per geometry: Σw/N within 3σ, and
    two-sample KS and χ² homogeneity tests (10 bins of equal probability under
    the oracle's sample) on the arrival time, the wavelength and the impact
    angle.  Thirty p-values are taken; under identical physics about a quarter
    of such runs would hold one below 0.01 by chance (two independent FP64
    oracle runs at the 30 m, 45° geometry give χ²(wavelength) p = 0.007).  So
    p > 0.01 is asserted per geometry family-wise: each of its six tests above
    0.01/6 (Bonferroni); and per test kind and observable, Fisher's combination
    over the five geometries above 0.01."""
    pv = {(k, o): [] for k in ("KS", "chi2") for o in OBSERVABLES}
    for dist, angle in GEOMS:
        a = math.radians(angle)
        tgt = pp.DetectionSphere((dist * math.sin(a), 0.0, dist * math.cos(a)), 0.21)
        n_gpu = 200_000_000
        g = ctx.propagate(cone_source(n_gpu), tgt, seed=1)
        rate = g.summary["n_hits"] / n_gpu
        n_cpu = int(min(max(3000 / max(rate, 1e-9), 1e6), 6e7))
        ch, cs = oracle_for(ctx, cone_source(n_cpu), tgt, seed=2, precision=64)
        say("geom", dist, angle, "gpu hits", g.summary["n_hits"], "cpu n", n_cpu, "cpu hits",
            cs["n_hits"])
        assert cs["n_hits"] > 150 and g.summary["n_hits"] > 150
        # per-photon weighted yield: the CPU's variance from its per-hit weights
        w = ch["weight"].astype(np.float64)
        mean_c = cs["sum_w"] / n_cpu
        sd_c = math.sqrt(max((w ** 2).sum() / n_cpu - mean_c ** 2, 0.0) / n_cpu)
        wg = g.hits["weight"].astype(np.float64)
        mean_g = g.summary["sum_w"] / n_gpu
        sd_g = math.sqrt(max((wg ** 2).sum() / n_gpu - mean_g ** 2, 0.0) / n_gpu)
        z = (mean_g - mean_c) / math.hypot(sd_c, sd_g)
        say("yield gpu %.4e cpu %.4e z=%.2f" % (mean_g, mean_c, z))
        assert abs(z) < 3.0

        def impact(h):
            n = (h["pos"].astype(np.float64) - np.array(tgt.position)) / tgt.radius
            return np.einsum("ij,ij->i", -h["dir"].astype(np.float64), n)

        for name, fg, fc in (("time", g.hits["time"], ch["time"]),
                             ("wavelength", g.hits["wavelength"], ch["wavelength"]),
                             ("impact", impact(g.hits), impact(ch))):
            fg = np.asarray(fg, np.float64)
            fc = np.asarray(fc, np.float64)
            p_ks = stats.ks_2samp(fg, fc).pvalue
            edges = np.unique(np.quantile(fc, np.linspace(0, 1, 11)))
            edges[0], edges[-1] = -np.inf, np.inf
            chi2, p_c, _, _ = stats.chi2_contingency(
                np.vstack([np.histogram(fg, edges)[0], np.histogram(fc, edges)[0]]))
            say("  %-10s KS p %.4f  chi2 %.2f p %.4f" % (name, p_ks, chi2, p_c))
            assert p_ks > 0.01 / 6 and p_c > 0.01 / 6, (dist, angle, name, p_ks, p_c)
            pv[("KS", name)].append(p_ks)
            pv[("chi2", name)].append(p_c)
    for key, ps in pv.items():
        pf = stats.chi2.sf(float(-2.0 * np.sum(np.log(ps))), 2 * len(ps))
        say("Fisher %s %s over %d geometries: p %.4f" % (key[0], key[1], len(ps), pf))
        assert pf > 0.01, (key, ps, pf)


# ------------------------------------------------------------- failure paths ----

def test_overflow_counts_exact(ctx):
    src = cone_source(300_000)
    tgt = pp.DetectionSphere((1.5, 0.0, 1.5), 0.3)
    full = ctx.propagate(src, tgt, seed=1)
    with pytest.raises(pp.PhotonPropError) as e:
        ctx.propagate(src, tgt, seed=1, hits_cap=5)
    assert e.value.code == pp._abi.PP_ERR_OVERFLOW
    part = ctx.propagate(src, tgt, seed=1, hits_cap=5, allow_overflow=True)
    assert part.summary["n_hits"] == full.summary["n_hits"]
    assert part.summary["sum_w"] == full.summary["sum_w"]  # sums are exact, bitwise
    assert part.summary["n_stored"] == 5 and part.hits.shape[0] == 5
    assert part.summary["status"] == pp._abi.PP_ERR_OVERFLOW
    assert set(part.hits["photon_id"].tolist()) <= set(full.hits["photon_id"].tolist())


def test_invalid_inputs_raise(ctx):
    tgt = pp.DetectionSphere((1.0, 0.0, 0.0), 0.3)
    with pytest.raises(pp.PhotonPropError, match="radius"):
        ctx.propagate(cone_source(10), pp.DetectionSphere((1.0, 0.0, 0.0), 0.0))
    with pytest.raises(pp.PhotonPropError, match="non-finite"):
        ctx.propagate(pp.PhotonSource((float("nan"), 0, 0), (0, 0, 1), 0.0, 10,
                                      pp.Monochromatic(400.0)), tgt)
    with pytest.raises(pp.PhotonPropError, match="Cherenkov"):
        ctx.propagate(cone_source(10, beta=0.5), tgt)
    with pytest.raises(pp.PhotonPropError, match="max_steps"):
        ctx.propagate(cone_source(10), tgt, max_steps=0)
    # a capacity whose byte size would wrap around 2^64 is refused before any
    # allocation or launch (raw call: the wrapper would allocate the buffer)
    import ctypes
    from photonprop.propagation import emitter_struct, spectrum_struct, sphere_struct
    spec, keep = spectrum_struct(pp.InverseSquareSpectrum(300.0, 600.0))
    small = np.zeros(4, dtype=pp._abi.HIT_DTYPE)
    summ = pp._abi.pp_summary()
    rc = ctx.lib.pp_propagate(ctx.handle, ctypes.byref(emitter_struct(cone_source(10))),
                              ctypes.byref(spec), ctypes.byref(sphere_struct(tgt)), 0, 0, 10,
                              pp._abi.ptr(small), 1 << 58, ctypes.byref(summ))
    assert rc == pp._abi.PP_ERR_INVALID_ARG and b"hits_cap" in ctx.lib.pp_last_error()
    with pytest.raises(pp.PhotonPropError, match="hit_pool_records"):
        pp.Server(ctx, pp.InverseSquareSpectrum(300.0, 600.0), hit_pool_records=1 << 40)
    # zero photons: an empty, exact result
    res = ctx.propagate(cone_source(0), tgt)
    assert res.summary["n_sim"] == 0 and res.summary["n_hits"] == 0 and res.hits.shape[0] == 0


def test_scale_geometries_statistics_vs_fp64_oracle(ctx):
    """The 64 bench geometries (16 distances geomspace(1, 100, 16) m x targets
    at 0/45/90/135 degrees from the emitter axis, r = 0.21 m): the batched kernel
    at 1e8 photons per geometry (seed 1, up to 20000 hit records kept per
    geometry) against the FP64 oracle at 1e7 (seed 2).  For every geometry with
    >= 200 oracle hits: the yield Σw/N within 4 sigma (the oracle's error from its
    hit sample, the GPU's from its hit count and the oracle's per-hit E[w²]) and
    two-sample KS tests with p > 1e-4; over all of them the
    yields' Σz² follows χ²(k) and Fisher's combination of the KS p-values
    χ²(2k), each with p > 0.001.  The KS tests cover the arrival time, the
    wavelength and the impact angle (cosine between the arrival direction and
    the inward surface normal).  (The arrival times are heavy-tailed, so their
    mean is compared through the whole distribution rather than a z-score.)"""
    from photonprop import workloads

    n_gpu, n_cpu, keep = 100_000_000, 10_000_000, 20_000
    spec = pp.InverseSquareSpectrum(300.0, 600.0)
    reqs = workloads.scale_geometries(photons=n_gpu, seed=1)
    gh, off, gs = ctx.propagate_batch(reqs, spec, per_request_cap=keep, hits_cap=len(reqs) * keep * 40,
                                      allow_overflow=True)
    assert np.all(gs["n_sim"] == n_gpu)
    zs, ps = [], {"time": [], "wavelength": [], "impact": []}
    for i, r in enumerate(reqs):
        src = cone_source(n_cpu)
        tgt = pp.DetectionSphere(tuple(r["center"]), float(r["radius"]))
        ch, cs = oracle_for(ctx, src, tgt, seed=2, stream=i, precision=64, hits_cap=2_000_000)
        if cs["n_hits"] < 200:
            continue
        assert cs["n_stored"] == cs["n_hits"]
        w = ch["weight"].astype(np.float64)
        ng = int(gs["n_hits"][i])
        yc, yg = cs["sum_w"] / n_cpu, gs["sum_w"][i] / n_gpu
        vc = ((w ** 2).sum() / n_cpu - yc ** 2) / n_cpu
        vg = ((w ** 2).mean() * ng / n_gpu - yg ** 2) / n_gpu
        zy = (yg - yc) / math.sqrt(vc + vg)
        seg = gh[int(off[i]):int(off[i + 1])]
        centre = np.asarray(r["center"], np.float64)

        def impact(h):
            nrm = (h["pos"].astype(np.float64) - centre) / float(r["radius"])
            return np.einsum("ij,ij->i", -h["dir"].astype(np.float64), nrm)

        pv = {}
        for name, fg, fc in (("time", seg["time"], ch["time"]),
                             ("wavelength", seg["wavelength"], ch["wavelength"]),
                             ("impact", impact(seg), impact(ch))):
            pv[name] = stats.ks_2samp(np.asarray(fg, np.float64), np.asarray(fc, np.float64)).pvalue
        say("geom %2d d=%6.2f hits gpu %8d (kept %5d) cpu %7d  yield z=%+.2f  KS p time %.3f "
            "wavelength %.3f impact %.3f" % (i, np.linalg.norm(r["center"]), ng, seg.shape[0],
                                            cs["n_hits"], zy, pv["time"], pv["wavelength"],
                                            pv["impact"]))
        assert abs(zy) < 4 and min(pv.values()) > 1e-4, (i, zy, pv)
        zs.append(zy)
        for name in ps:
            ps[name].append(pv[name])
    k = len(zs)
    p_y = stats.chi2.sf(float(np.sum(np.square(zs))), k)
    p_f = {name: stats.chi2.sf(float(-2.0 * np.sum(np.log(v))), 2 * k) for name, v in ps.items()}
    say("geometries compared %d: yield chi2 p %.3f, Fisher KS p %s" % (k, p_y, p_f))
    assert k >= 20 and p_y > 0.001 and min(p_f.values()) > 0.001
