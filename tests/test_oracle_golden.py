"""Pins of the CPU oracle: known answers from the reference's unit tests
(src/tests/test_photon_prob.jl), Philox vectors from rocRAND's header,
closed forms and brute force.  CPU only."""
import json
import math
import os

import numpy as np
import pytest
from scipy import stats

import oracle
import photonprop as pp
from photonprop import _abi
from photonprop.propagation import emitter_struct, spectrum_struct, sphere_struct

from closed_forms import C_LIGHT, beta_below, check_unscattered_hits

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------- Philox ----

def test_philox_known_answers():
    kat = json.load(open(os.path.join(HERE, "golden", "philox_kat.json")))
    for case in kat["cases"]:
        out = oracle.philox4x32_10(case["counter"], case["key"])
        assert list(out) == case["out"], case


def test_philox_published_vectors():
    # Random123 kat_vectors for philox4x32_10 (Salmon et al. 2011).
    assert list(oracle.philox4x32_10([0, 0, 0, 0], [0, 0])) == [
        0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(oracle.philox4x32_10([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]


def test_uniform_range_and_ends():
    assert oracle.uniform(0) == 2.0 ** -24
    assert oracle.uniform(0xFFFFFFFF) == 1.0
    assert oracle.uniform(0x80000000) == 0.5 + 2.0 ** -24
    rng = np.random.default_rng(0)
    for w in rng.integers(0, 2 ** 32, 1000, dtype=np.uint64):
        u = oracle.uniform(int(w))
        assert 0.0 < u <= 1.0
        assert u == ((int(w) >> 8) + 1) * 2.0 ** -24


# --------------------------------------------------------------- uniform ----

def test_block_draws_range_uniformity_and_independence():
    """Per Philox block the propagation draws four 23-bit uniforms on (0, 1]
    (the words' top 23 bits) and two 16-bit azimuths on [0, 1) (the low bytes
    of word pairs): each on its lattice, uniform (KS), and the six mutually
    uncorrelated (bits 8 of the words are the only ones left unused)."""
    blocks = [oracle.philox4x32_10([i, 7, 0, 3], [11, 0]) for i in range(20000)]
    d = [oracle.block_draws(b) for b in blocks]
    u = np.array([x[0] for x in d], dtype=np.float64)
    p = np.array([x[1] for x in d], dtype=np.float64)
    assert u.min() > 0.0 and u.max() <= 1.0 and p.min() >= 0.0 and p.max() < 1.0
    assert np.all(u * 2 ** 23 == np.round(u * 2 ** 23))
    assert np.all(p * 2 ** 16 == np.round(p * 2 ** 16))
    allv = np.concatenate([u, p], axis=1)
    for j in range(6):
        assert stats.kstest(allv[:, j], "uniform").pvalue > 0.001, j
    c = np.corrcoef(allv.T)
    assert np.abs(c - np.eye(6)).max() < 0.04
    # lattice ends: word 0 -> u = 1 (the device's a = 1), all ones -> u = 2^-23
    u0, p0 = oracle.block_draws([0, 0, 0, 0])
    np.testing.assert_array_equal(u0, [1.0] * 4)
    np.testing.assert_array_equal(p0, [0.0, 0.0])
    u1, p1 = oracle.block_draws([0xFFFFFFFF] * 4)
    np.testing.assert_array_equal(u1, [2.0 ** -23] * 4)
    np.testing.assert_array_equal(p1, [1.0 - 2.0 ** -16] * 2)
    # u_k uses bits 9..31 only; p0 = byte 0 of word 0 (high) and of word 1 (low)
    assert oracle.block_draws([0x1FF, 0, 0, 0])[0][0] == 1.0
    assert oracle.block_draws([0x200, 0, 0, 0])[0][0] == 1.0 - 2.0 ** -23
    assert oracle.block_draws([0x01, 0x00, 0, 0])[1][0] == 2.0 ** -8
    assert oracle.block_draws([0x00, 0x01, 0, 0])[1][0] == 2.0 ** -16
    assert oracle.block_draws([0x100, 0x100, 0, 0])[1][0] == 0.0  # bit 8 unused

def test_uniform_range_reference_kat():
    # src/tests/test_photon_prob.jl:13-17: uniform(2, 5) stays in [2, 5].
    words = [oracle.philox4x32_10([i, 0, 0, 0], [0, 0]) for i in range(2500)]
    u = np.array([oracle.uniform(int(w)) for ws in words for w in ws])
    v = 2.0 + u * (5.0 - 2.0)
    assert v.min() >= 2.0 and v.max() <= 5.0


# -------------------------------------------------------------------- HG ----

def hg_cdf(mu, g):
    """Closed-form CDF of the Henyey-Greenstein phase function in cosθ."""
    return (1 - g * g) / (2 * g) * (1 / np.sqrt(1 + g * g - 2 * g * mu) - 1 / (1 + g))


@pytest.mark.parametrize("g", [0.92, 0.5, 0.99, -0.3])
def test_hg_inverse_cdf_is_inverse_of_closed_form(g):
    eta = np.linspace(1e-4, 1 - 1e-4, 2001)
    mu = np.array([oracle.hg_costheta(g, float(e), precision=64) for e in eta])
    np.testing.assert_allclose(hg_cdf(mu, g), eta, atol=1e-9)
    mu32 = np.array([oracle.hg_costheta(g, float(e), precision=32) for e in eta])
    np.testing.assert_allclose(mu32, mu, atol=2e-5)


@pytest.mark.parametrize("g", [0.92, 0.97, 0.0])
def test_hg_moments_quadrature(g):
    # Midpoint rule over η in (0, 1): E[cosθ] = g, E[cos²θ] = (1 + 2g²)/3.
    n = 200_000
    eta = (np.arange(n) + 0.5) / n
    mu = np.array([oracle.hg_costheta(g, float(e), precision=64) for e in eta[::10]])
    # coarse grid for speed, then a full-resolution check of the mean
    assert abs(mu.mean() - g) < 2e-3
    assert abs((mu ** 2).mean() - (1 + 2 * g * g) / 3) < 2e-3


def test_hg_mean_reference_kat():
    # src/tests/test_photon_prob.jl:19-21: mean cosθ of 1e5 draws at g=0.97 ≈ g (atol 0.01).
    g = 0.97
    ws = np.concatenate([oracle.philox4x32_10([i, 0, 0, 0], [0, 0]) for i in range(25_000)])
    mu = np.array([oracle.hg_costheta(g, oracle.uniform(int(w))) for w in ws])
    assert abs(mu.mean() - g) < 0.01


# -------------------------------------------------------------- rotation ----

def test_rotate_to_axis_reference_kats():
    ey = np.array([0.0, 1.0, 0.0])
    # src/tests/test_photon_prob.jl:26-30
    np.testing.assert_array_equal(oracle.rotate_to_axis([0, 0, 1], ey, 64), ey)
    np.testing.assert_array_equal(oracle.rotate_to_axis([0, 0, -1], ey, 64), -ey)
    # :37-41 — rotating e_z to ŷ maps ŷ to -e_z
    np.testing.assert_allclose(oracle.rotate_to_axis(ey, ey, 64), [0, 0, -1], atol=1e-15)
    # :43-53 — unit norm for a generic pair
    a = np.array([0.2, 0.3, 1.0]); a /= np.linalg.norm(a)
    b = np.array([-5.0, 1.0, -20.0]); b /= np.linalg.norm(b)
    assert abs(np.linalg.norm(oracle.rotate_to_axis(a, b, 64)) - 1) < 1e-14
    assert abs(np.linalg.norm(oracle.rotate_to_axis(a, b, 32)) - 1) < 1e-6


def test_rotate_to_axis_equals_rodrigues():
    # src/tests/test_photon_prob.jl:55-64 against rodrigues_rotation (:161-169).
    rng = np.random.default_rng(1)
    ez = np.array([0.0, 0.0, 1.0])
    for _ in range(500):
        a = rng.normal(size=3); a /= np.linalg.norm(a)
        v = rng.normal(size=3); v /= np.linalg.norm(v)
        r64 = oracle.rotate_to_axis(a, v, 64)
        rr = oracle.rodrigues_rotation(ez, a, v)
        np.testing.assert_allclose(r64, rr, atol=1e-12)
        # FP32 against FP64 on the same FP32-rounded inputs; the reference's form
        # loses accuracy as old_dir approaches -e_z (1 - z² cancels), so only
        # directions away from the south pole are held to 3e-6.
        a32, v32 = a.astype(np.float32), v.astype(np.float32)
        if a32[2] > -0.9:
            rr32 = oracle.rodrigues_rotation(ez, a32.astype(np.float64) / np.linalg.norm(a32.astype(np.float64)),
                                             v32.astype(np.float64))
            np.testing.assert_allclose(oracle.rotate_to_axis(a32, v32, 32), rr32, atol=3e-6)
        # invariant: the angle to the old direction equals the local polar angle
        assert abs(np.dot(r64, a) - v[2]) < 1e-12


def test_rodrigues_is_a_rotation_matrix():
    # Rodrigues (textbook) against an explicit rotation matrix from its axis/angle.
    rng = np.random.default_rng(2)
    ez = np.array([0.0, 0.0, 1.0])
    for _ in range(50):
        a = rng.normal(size=3); a /= np.linalg.norm(a)
        k = np.cross(ez, a); s = np.linalg.norm(k); k /= s; c = a[2]
        K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
        R = np.eye(3) + s * K + (1 - c) * K @ K
        np.testing.assert_allclose(R @ ez, a, atol=1e-12)
        v = rng.normal(size=3)
        np.testing.assert_allclose(oracle.rodrigues_rotation(ez, a, v), R @ v, atol=1e-12)


def test_update_position_kat():
    # src/tests/test_photon_prob.jl:67-73
    np.testing.assert_allclose(oracle.update_position([0, 0, 0], [0, 1, 0], 10.0, 64), [0, 10, 0])


# --------------------------------------------------------- interpolation ----

@pytest.mark.parametrize("precision", [32, 64])
def test_interpolation_kats(precision):
    # src/tests/test_photon_prob.jl:75-90 with f(x) = x² on 0:4
    xs = np.arange(5.0)
    ys = xs ** 2
    for x, want in ((0.5, 0.5), (1.5, 2.5), (4.5, 16.0), (-1.0, 0.0)):
        assert oracle.fast_linear_interp_sorted(x, xs, ys, precision) == pytest.approx(want)
        assert oracle.fast_linear_interp_knots(x, ys, 0.0, 4.0, precision) == pytest.approx(want)
        assert pp.utils.fast_linear_interp(x, xs, ys) == pytest.approx(want)
        assert pp.utils.fast_linear_interp_knots(x, ys, 0.0, 4.0) == pytest.approx(want)


# ------------------------------------------------------------ ray–sphere ----

def brute_force_entry(pos, d, c, r, step, n=200_001):
    t = np.linspace(0, step, n)
    inside = np.linalg.norm(pos[None, :] + t[:, None] * d[None, :] - c[None, :], axis=1) <= r
    if inside[0]:
        return None  # starts inside: the test is about entering
    idx = np.nonzero(inside)[0]
    return None if idx.size == 0 else t[idx[0]]


def test_ray_sphere_special_cases():
    c = np.array([10.0, 0.0, 0.0])
    # head-on: entry at 10 - 0.21
    hit, d = oracle.ray_sphere([0, 0, 0], [1, 0, 0], c, 0.21, 20.0, 64)
    assert hit and d == pytest.approx(9.79)
    # beyond the step: no hit
    assert not oracle.ray_sphere([0, 0, 0], [1, 0, 0], c, 0.21, 9.0, 64)[0]
    # behind (D4): no hit
    assert not oracle.ray_sphere([0, 0, 0], [-1, 0, 0], c, 0.21, 50.0, 64)[0]
    # inside the sphere: entry distance negative, no hit
    assert not oracle.ray_sphere([10, 0, 0], [1, 0, 0], c, 0.21, 5.0, 64)[0]
    # miss by a hair
    assert not oracle.ray_sphere([0, 0.2101, 0], [1, 0, 0], c, 0.21, 20.0, 64)[0]
    assert oracle.ray_sphere([0, 0.2099, 0], [1, 0, 0], c, 0.21, 20.0, 64)[0]


def test_ray_sphere_brute_force():
    rng = np.random.default_rng(3)
    c = np.array([1.0, -0.5, 2.0])
    for _ in range(300):
        pos = rng.uniform(-3, 3, size=3)
        if np.linalg.norm(pos - c) <= 0.5:
            continue
        tgt = c + rng.normal(scale=0.4, size=3)
        d = tgt - pos; d /= np.linalg.norm(d)
        step = rng.uniform(0.1, 8.0)
        hit, dist = oracle.ray_sphere(pos, d, c, 0.5, step, 64)
        bf = brute_force_entry(pos, d, c, 0.5, step)
        if bf is None:
            assert not hit or dist > step - 1e-4
        else:
            assert hit and abs(dist - bf) < 1e-4


# ---------------------------------------------------------- emission ----

def _structs(medium, spectrum, profile, axis=(0.0, 0.0, 1.0), n=1000, beta=1.0, t0=0.0):
    ctx_tables = pp.medium_tables(medium)
    keep = [np.ascontiguousarray(a, dtype=np.float32)
            for a in (ctx_tables.n_phase, ctx_tables.abs_len, ctx_tables.sca_len)]
    t = _abi.pp_medium_tables()
    t.lambda_min, t.lambda_max, t.n_points = ctx_tables.lambda_min, ctx_tables.lambda_max, ctx_tables.n_points
    t.n_phase, t.abs_len, t.sca_len = [a.ctypes.data_as(_abi.c_float_p) for a in keep]
    t.g = ctx_tables.g
    src = pp.PhotonSource((0, 0, 0), axis, t0, n, spectrum, pp.AngularEmissionProfile(profile), beta)
    sp, k2 = spectrum_struct(spectrum)
    return t, sp, emitter_struct(src), (keep, k2, ctx_tables)


def test_wavelength_inverse_square_closed_form():
    m = pp.SeawaterConfigMedium()
    t, sp, em, keep = _structs(m, pp.InverseSquareSpectrum(300, 600), "CherenkovCone")
    out = oracle.emit(t, sp, em, 0, 0, np.arange(4000), precision=64)
    lam = out[:, 0]
    u = np.array([oracle.block_draws(oracle.philox4x32_10([0, i, 0, 0], [0, 0]))[0][0]
                  for i in range(4000)], dtype=np.float64)
    np.testing.assert_allclose(lam, 600.0 / (2.0 - u), rtol=1e-12)
    # KS against the closed-form CDF, and the mean 600·ln 2
    cdf = pp.InverseSquareSpectrum(300, 600).cdf
    assert stats.kstest(lam, cdf).pvalue > 0.01
    # mean 600·ln 2; Var λ = 600·300 - (600 ln 2)² (density 600/λ² on [300, 600]),
    # so the mean of 4000 draws has sigma 83.8/sqrt(4000) = 1.33 nm: 4 sigma
    sigma = math.sqrt(600.0 * 300.0 - (600 * math.log(2)) ** 2) / math.sqrt(lam.size)
    assert abs(lam.mean() - 600 * math.log(2)) < 4 * sigma


def test_cherenkov_cone_angle_and_azimuth():
    m = pp.SeawaterConfigMedium()
    axis = np.array([0.3, -0.4, 0.866]); axis /= np.linalg.norm(axis)
    t, sp, em, keep = _structs(m, pp.InverseSquareSpectrum(300, 600), "CherenkovCone", tuple(axis))
    out = oracle.emit(t, sp, em, 5, 3, np.arange(3000), precision=32)
    lam, d = out[:, 0].astype(np.float64), out[:, 1:].astype(np.float64)
    n = m.refractive_index(lam)
    np.testing.assert_allclose(d @ axis, 1.0 / n, atol=2e-6)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    # azimuth about the axis is uniform
    e1 = np.cross(axis, [1.0, 0, 0]); e1 /= np.linalg.norm(e1); e2 = np.cross(axis, e1)
    phi = (np.arctan2(d @ e2, d @ e1) / (2 * np.pi)) % 1.0
    assert stats.kstest(phi, "uniform").pvalue > 0.01


def test_reference_isotropic_profile():
    # θ ~ U(-1, 1) rad, φ ~ U(0, 2π), axis ignored (src/photon_prop_cuda.jl:52-56).
    m = pp.SeawaterConfigMedium()
    t, sp, em, keep = _structs(m, pp.Monochromatic(450.0), "IsotropicEmission", (1.0, 0.0, 0.0))
    out = oracle.emit(t, sp, em, 0, 0, np.arange(3000), precision=64)
    ct = out[:, 3]
    assert ct.min() >= math.cos(1.0) - 1e-12
    # cosθ = cos(U(-1,1)) has CDF P(cos θ ≤ c) = 1 - acos(c)
    assert stats.kstest(ct, lambda c: 1 - np.arccos(np.clip(c, -1, 1))).pvalue > 0.01


# ------------------------------------------------- propagation closed forms ----

class _Med(pp.medium.MediumProperties):
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


def _run_oracle(medium, source, target, max_steps=10, seed=0, precision=32, n_threads=4):
    from photonprop._abi import HIT_DTYPE, pp_summary
    t, sp, em, keep = _structs(medium, source.spectrum, source.emission_profile.kind,
                               source.direction, source.photons, source.beta, source.time)
    hits, summ = oracle.propagate(t, sp, em, sphere_struct(target), seed=seed, max_steps=max_steps,
                                  precision=precision, n_threads=n_threads, hit_dtype=HIT_DTYPE,
                                  summary_struct=pp_summary)
    return hits, summ


def test_no_scatter_cone_hit_fraction_closed_form():
    # λ_sca → ∞: the cone ring meets a sphere of angular radius ρ = asin(r/d) at
    # angle α from the axis for a fraction acos(x)/π of azimuths,
    # x = (cos ρ - cos θc cos α)/(sin θc sin α).
    n, d, r = 1.34, 5.0, 0.4
    thc = math.acos(1 / n)
    alpha = thc + 0.03
    med = _Med(n=n)
    src = pp.PhotonSource((0, 0, 0), (0, 0, 1), 0.0, 200_000, pp.Monochromatic(450.0),
                          pp.AngularEmissionProfile("CherenkovCone"))
    tgt = pp.DetectionSphere((d * math.sin(alpha), 0.0, d * math.cos(alpha)), r)
    hits, summ = _run_oracle(med, src, tgt)
    rho = math.asin(r / d)
    x = (math.cos(rho) - math.cos(thc) * math.cos(alpha)) / (math.sin(thc) * math.sin(alpha))
    p = math.acos(x) / math.pi
    N = src.photons
    assert abs(summ.n_hits - N * p) < 5 * math.sqrt(N * p * (1 - p)), (summ.n_hits, N * p)
    # a hit ends on its first step; a miss runs all 10 steps
    assert summ.n_steps == summ.n_hits + 10 * (N - summ.n_hits)
    # λ_abs finite: every weight is exp(-L/λ_abs) for its own path length
    np.testing.assert_allclose(hits["weight"], np.exp(-hits["dist"] / 20.0), rtol=1e-6)


def test_straight_line_survival_exp_minus_d_over_abs():
    # n·β = 1 makes the cone collapse onto the axis: every photon flies straight
    # into a sphere on the axis, so Σw/N = exp(-(d - r)/λ_abs) (to 1e-3).
    n, d, r, la = 1.34, 12.0, 0.21, 25.0
    med = _Med(n=n, abs_len=la)
    src = pp.PhotonSource((0, 0, 0), (0, 0, 1), 0.0, 20_000, pp.Monochromatic(450.0),
                          pp.AngularEmissionProfile("CherenkovCone"), beta=1.0 / n)
    tgt = pp.DetectionSphere((0.0, 0.0, d), r)
    hits, summ = _run_oracle(med, src, tgt)
    assert summ.n_hits == src.photons
    assert abs(summ.sum_w / src.photons - math.exp(-(d - r) / la)) < 1e-3


@pytest.mark.parametrize("d", [0.8, 2.3, 6.0, 15.3])
def test_unscattered_fraction_exp_minus_d_over_sca(d):
    # With one step allowed, a photon on the axis reaches the sphere iff its first
    # exponential step exceeds d - r: P = exp(-(d - r)/λ_sca).  The four distances
    # pin the step-length CDF at P = 0.94, 0.76, 0.49 and 0.15.
    n, r, ls = 1.34, 0.3, 8.0
    med = _Med(n=n, abs_len=1e30, sca_len=ls)
    src = pp.PhotonSource((0, 0, 0), (0, 0, 1), 0.0, 100_000, pp.Monochromatic(450.0),
                          pp.AngularEmissionProfile("CherenkovCone"), beta=1.0 / n)
    tgt = pp.DetectionSphere((0.0, 0.0, d), r)
    hits, summ = _run_oracle(med, src, tgt, max_steps=1)
    p = math.exp(-(d - r) / ls)
    N = src.photons
    assert abs(summ.n_hits - N * p) < 5 * math.sqrt(N * p * (1 - p))
    assert np.all(hits["weight"] == 1.0)  # λ_abs → ∞ gives w ≡ 1
    assert summ.n_steps == N


def test_fp32_and_fp64_oracles_agree_per_photon():
    m = pp.SeawaterConfigMedium()
    src = pp.PhotonSource((0, 0, 0), (0, 0, 1), 0.0, 100_000, pp.InverseSquareSpectrum(300, 600),
                          pp.AngularEmissionProfile("CherenkovCone"))
    tgt = pp.DetectionSphere((1.0, 0.0, 1.2), 0.4)
    h32, s32 = _run_oracle(m, src, tgt, precision=32)
    h64, s64 = _run_oracle(m, src, tgt, precision=64)
    common = np.intersect1d(h32["photon_id"], h64["photon_id"])
    assert s32.n_hits > 500
    assert len(common) >= s64.n_hits - max(3, 1e-3 * s64.n_hits)
    assert abs(s32.sum_w - s64.sum_w) < 1e-3 * s64.sum_w


# --------------------------------------------- hit time and FP64 moments ----

@pytest.mark.parametrize("precision", [32, 64])
def test_straight_line_hit_time_and_time_moments(precision):
    """Photons emitted straight along the axis into a sphere on the axis
    (constant n, no scattering): every hit has L = d - r, time
    t0 + (d - r)·n/c (the phase index, src/photon_prop_cuda.jl:634-638) and
    weight exp(-(d - r)/λ_abs), so Σw·t = t·Σw and Σw·t² = t²·Σw with that one t."""
    n, d, r, la, t0 = 1.34, 12.0, 0.21, 25.0, 17.25
    med = _Med(n=n, abs_len=la)
    src = pp.PhotonSource((0, 0, 0), (0, 0, 1), t0, 5_000, pp.Monochromatic(450.0),
                          pp.AngularEmissionProfile("CherenkovCone"), beta=beta_below(n))
    hits, summ = _run_oracle(med, src, pp.DetectionSphere((0.0, 0.0, d), r), precision=precision)
    L = d - r
    t = t0 + L * n / C_LIGHT
    w = math.exp(-L / la)
    assert summ.n_hits == src.photons == hits.shape[0]
    np.testing.assert_allclose(hits["dist"], L, rtol=1e-6)
    np.testing.assert_allclose(hits["time"], t, rtol=1e-6)
    np.testing.assert_allclose(hits["weight"], w, rtol=1e-6)
    assert summ.sum_w == pytest.approx(src.photons * w, rel=1e-6)
    assert summ.sum_wt == pytest.approx(summ.sum_w * t, rel=1e-6)
    assert summ.sum_wt2 == pytest.approx(summ.sum_w * t * t, rel=1e-6)


def test_unscattered_hit_time_weight_and_moments_closed_form():
    """Uniform emission in the configured seawater with one step allowed: the
    hits are the unscattered photons, whose time, weight and path length follow
    from the geometry and the medium's formulas alone (a wavelength-dependent
    n, so a misplaced or mis-indexed n fails)."""
    m = pp.SeawaterConfigMedium()
    centre, r, t0 = (1.1, -0.4, 2.3), 0.5, 3.75
    src = pp.PhotonSource((0, 0, 0), (0, 0, 1), t0, 200_000, pp.InverseSquareSpectrum(300, 600),
                          pp.AngularEmissionProfile("UniformSphere"))
    hits, summ = _run_oracle(m, src, pp.DetectionSphere(centre, r), max_steps=1)
    s = {k: getattr(summ, k) for k in ("n_stored", "n_hits", "sum_w", "sum_wt", "sum_wt2")}
    assert s["n_hits"] > 1000
    check_unscattered_hits(hits, s, (0, 0, 0), centre, r, t0)
