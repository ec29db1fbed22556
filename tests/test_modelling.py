"""Result-table post-processing (src/modelling.jl:40-111, src/emission.jl:32-99,
src/detection.jl:18-33): pinned by quadrature, scipy's Gamma MLE, the
reference's PMT acceptance data and structural properties.  CPU only except
the last test."""
import math
import os

import numpy as np
import pandas as pd
import pytest
from scipy import integrate, stats

import photonprop as pp
from photonprop import modelling as md
from photonprop.detection import PMTWavelengthAcceptance, p_one_pmt_acc
from photonprop.emission import cherenkov_ang_dist

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("n", [1.1, 1.2, 1.33, 1.41, 1.5])
@pytest.mark.parametrize("lo,hi", [(-1.0, 1.0), (0.2, 0.9), (-0.5, 0.3)])
def test_angular_integral_closed_form_vs_quadrature(n, lo, hi):
    want, _ = integrate.quad(lambda x: cherenkov_ang_dist(x, n), lo, hi, points=[1 / n], limit=200,
                             epsabs=1e-12, epsrel=1e-10)
    assert md.cherenkov_ang_dist_integral(n, lo, hi) == pytest.approx(want, rel=1e-8)


def test_angular_integral_table_interpolation():
    for n in (1.1, 1.25, 1.5):
        assert md.cherenkov_ang_dist_int(n) == pytest.approx(md.cherenkov_ang_dist_integral(n), rel=1e-12)
    mid = md.cherenkov_ang_dist_int(1.335)
    assert min(md.cherenkov_ang_dist_int(1.33), md.cherenkov_ang_dist_int(1.34)) <= mid
    assert mid <= max(md.cherenkov_ang_dist_int(1.33), md.cherenkov_ang_dist_int(1.34))
    with pytest.raises(ValueError):
        md.cherenkov_ang_dist_int(1.6)


def test_dir_reweight_peaks_on_the_cherenkov_angle():
    n = 1.34
    th = np.linspace(0, math.pi, 20001)
    w = md.get_dir_reweight(th, 0.0, np.full_like(th, n))
    assert abs(th[np.argmax(w)] - math.acos(1 / n)) < 1e-3
    # shifting the observation angle shifts the weight curve
    w2 = md.get_dir_reweight(th, 0.3, np.full_like(th, n))
    assert abs(th[np.argmax(w2)] - (math.acos(1 / n) + 0.3)) < 1e-3
    # normalisation: ∫ f(cosθ) dcosθ / (2∫) = 1/2
    c = np.linspace(-1, 1, 200001)
    f = cherenkov_ang_dist(c, n) / (2 * md.cherenkov_ang_dist_int(n))
    assert np.trapezoid(f, c) == pytest.approx(0.5, rel=1e-4)


def test_weighted_gamma_mle():
    rng = np.random.default_rng(0)
    x = rng.gamma(2.5, 7.0, 200_000)
    a, th = md.fit_gamma_weighted(x, np.ones_like(x))
    a_sp, _, th_sp = stats.gamma.fit(x, floc=0)
    assert a == pytest.approx(a_sp, rel=1e-4) and th == pytest.approx(th_sp, rel=1e-4)
    assert a == pytest.approx(2.5, rel=2e-2) and th == pytest.approx(7.0, rel=2e-2)
    # integer weights are frequency weights
    xs = rng.gamma(1.7, 3.0, 500)
    k = rng.integers(1, 5, 500)
    aw, tw = md.fit_gamma_weighted(xs, k.astype(float))
    ar, tr = md.fit_gamma_weighted(np.repeat(xs, k), np.ones(k.sum()))
    assert aw == pytest.approx(ar, rel=1e-10) and tw == pytest.approx(tr, rel=1e-10)


def test_pmt_acceptance_table():
    d = np.loadtxt(os.path.join(HERE, "golden", "pmt_acceptance.csv"), delimiter=",")
    np.testing.assert_allclose(p_one_pmt_acc(d[:, 0]), d[:, 1], rtol=0, atol=0)
    mid = 0.5 * (d[10, 0] + d[11, 0])
    assert p_one_pmt_acc(mid) == pytest.approx(0.5 * (d[10, 1] + d[11, 1]))
    assert np.all((d[:, 1] >= 0) & (d[:, 1] <= 1))
    with pytest.raises(ValueError):
        p_one_pmt_acc(150.0)
    with pytest.raises(ValueError):
        PMTWavelengthAcceptance([1, 1], [0, 0])


def test_sobol_points():
    np.testing.assert_allclose(md.sobol_1d(7, 0.0, 1.0),
                               [0.5, 0.75, 0.25, 0.375, 0.875, 0.625, 0.125])
    np.testing.assert_allclose(md.sobol_1d(3, 0.0, math.pi), [math.pi / 2, 0.75 * math.pi, math.pi / 4])


def test_fit_photon_dist_on_a_synthetic_table():
    rng = np.random.default_rng(1)
    n = 20000
    df = pd.DataFrame({
        "tres": rng.gamma(2.0, 5.0, n) - 0.5,
        "initial_theta": rng.uniform(0, 1.0, n),
        "ref_ix": np.full(n, 1.34),
        "abs_weight": rng.uniform(0.2, 1.0, n),
        "dist_travelled": rng.uniform(5, 50, n),
        "wavelength": rng.uniform(300, 600, n),
    })
    fits = md.fit_photon_dist([0.2, 0.8], df, 10 ** 6)
    assert list(fits.columns) == ["obs_angle", "fit_alpha", "fit_theta", "det_fraction"]
    tot = (md.get_dir_reweight(df["initial_theta"], 0.8, df["ref_ix"]) * df["abs_weight"]
           * p_one_pmt_acc(df["wavelength"]))
    assert fits["det_fraction"][1] == pytest.approx(tot.sum() / 1e6, rel=1e-12)
    assert np.all(fits["fit_alpha"] > 0) and np.all(fits["fit_theta"] > 0)


@pytest.mark.gpu
def test_make_photon_fits_end_to_end():
    """make_photon_fits on the GPU (src/modelling.jl:86-111); one cell's
    det_fraction is recomputed by hand from the raw hit records of the same
    propagation (the reference's setup: cascadia water, 20-knot spectrum,
    :IsotropicEmission, sphere r = 0.21 m on the z axis, seed 0):
    Σ f(cos(θ0 - obs))/(2∫f) · w · PMT(λ) / N with θ0 the initial polar angle."""
    n_ph = 2_000_000
    fits = md.make_photon_fits(n_ph, 3, 4, max_dist=30.0)
    assert set(fits.columns) == {"obs_angle", "fit_alpha", "fit_theta", "det_fraction", "distance"}
    assert len(fits) > 0 and np.all(np.isfinite(fits["fit_alpha"]))
    assert np.all(fits["det_fraction"] >= 0)
    row = fits.iloc[len(fits) - 1]
    m = pp.make_cascadia_medium_properties()
    src = pp.PhotonSource((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0.0, n_ph,
                          pp.CherenkovSpectrum((300.0, 800.0), 20, m),
                          pp.AngularEmissionProfile("IsotropicEmission"))
    with pp.Context(m) as c:
        h = c.propagate(src, pp.DetectionSphere((0.0, 0.0, float(row["distance"])), 0.21), seed=0).hits
    theta0 = np.arccos(np.clip(h["init_dir"][:, 2].astype(np.float64), -1.0, 1.0))
    lam = h["wavelength"].astype(np.float64)
    n = pp.medium.get_refractive_index_fry(lam, m.salinity, m.temperature, m.pressure)
    f = cherenkov_ang_dist(np.cos(theta0 - float(row["obs_angle"])), n)
    norm = 2.0 * md.cherenkov_ang_dist_integral(n)
    hand = float((f / norm * h["weight"].astype(np.float64) * p_one_pmt_acc(lam)).sum() / n_ph)
    print("det_fraction", float(row["det_fraction"]), "by hand", hand, flush=True)
    assert hand > 0
    # the table interpolates the angular integral over n in steps of 0.01
    assert float(row["det_fraction"]) == pytest.approx(hand, rel=1e-4)
