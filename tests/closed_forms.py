"""Closed forms shared by the oracle's and the library's tests: the hit time
and weight of photons whose path follows from geometry alone, in the
configured seawater (n linear from 1.36 at 300 nm to 1.34 at 600 nm; λ_abs
30 m at 450 nm falling linearly to 10 m at 300 and 600 nm)."""
import math

import numpy as np
import pytest

C_LIGHT = 0.299792458  # m/ns, exact (SI)


def beta_below(n):
    """The largest FP32 speed below 1/n by two ulps: n·β < 1, so the Cherenkov
    cone collapses onto the emitter axis (cosθ = 1/(nβ) > 1 is clamped to 1)."""
    b = np.float32(1.0 / np.float32(n))
    return float(np.nextafter(np.nextafter(b, np.float32(0)), np.float32(0)))


def config_n(lam):
    """n(λ) of the configured seawater: linear from 1.36 at 300 nm to 1.34 at 600 nm."""
    return 1.36 - 0.02 * (np.asarray(lam, np.float64) - 300.0) / 300.0


def config_abs(lam):
    """λ_abs of the configured seawater: 30 m at 450 nm falling linearly to 10 m at
    300 and 600 nm."""
    return 30.0 - 20.0 * np.abs(np.asarray(lam, np.float64) - 450.0) / 150.0


def sphere_entry(origin, direction, centre, r):
    """FP64 entry distance along unit rays, and the impact parameter b."""
    d = direction / np.linalg.norm(direction, axis=1, keepdims=True)
    q = np.asarray(origin, np.float64)[None, :] - np.asarray(centre, np.float64)[None, :]
    a = (d * q).sum(1)
    b2 = (q * q).sum(1) - a * a
    return -a - np.sqrt(np.maximum(r * r - b2, 0.0)), np.sqrt(np.maximum(b2, 0.0))


def check_unscattered_hits(hits, summ, origin, centre, r, t0, max_grazing=0.01):
    """Closed-form checks of unscattered hits in the configured seawater (shared
    by the oracle and the GPU tests): each record's time is t0 + L·n(λ)/c and
    its weight exp(-L/λ_abs(λ)) with n and λ_abs the config's formulas at the
    record's wavelength; L is the FP64 entry distance along the recorded
    emission direction (held to the FP32 error that the ray's impact parameter
    b allows: dL/db = b/sqrt(r² - b²)); the summary's FP64 moments are the
    sums of w, w·t and w·t² over the records."""
    lam = hits["wavelength"].astype(np.float64)
    L = hits["dist"].astype(np.float64)
    t = hits["time"].astype(np.float64)
    w = hits["weight"].astype(np.float64)
    np.testing.assert_allclose(t, t0 + L * config_n(lam) / C_LIGHT, rtol=2e-6)
    np.testing.assert_allclose(w, np.exp(-L / config_abs(lam)), rtol=1e-5)
    L64, b = sphere_entry(origin, hits["init_dir"].astype(np.float64), centre, r)
    slack = r * r - b * b
    ok = slack > 1e-6 * r * r
    assert (~ok).mean() <= max_grazing
    tol = 1e-5 * L64 + 1e-6 * L64 * b / np.sqrt(np.where(ok, slack, 1.0)) + 1e-5
    assert np.all(np.abs(L - L64)[ok] <= tol[ok]), np.abs(L - L64)[ok].max()
    assert summ["n_stored"] == summ["n_hits"] == hits.shape[0]
    assert summ["sum_w"] == pytest.approx(w.sum(), rel=1e-10)
    assert summ["sum_wt"] == pytest.approx((w * t).sum(), rel=1e-10)
    assert summ["sum_wt2"] == pytest.approx((w * t * t).sum(), rel=1e-10)
