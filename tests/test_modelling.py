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
    fits = md.make_photon_fits(2_000_000, 3, 4, max_dist=30.0)
    assert set(fits.columns) == {"obs_angle", "fit_alpha", "fit_theta", "det_fraction", "distance"}
    assert len(fits) > 0 and np.all(np.isfinite(fits["fit_alpha"]))
    assert np.all(fits["det_fraction"] >= 0)
