"""Result-table post-processing of PhotonProp.jl's Modelling module
(src/modelling.jl:40-111): Cherenkov angular re-weighting of the propagated
photons, PMT acceptance weighting, weighted Gamma fits of the residual time
per observation angle, and the table builder make_photon_fits that drives them
over Sobol-spaced distances with the GPU propagate_distance.
"""
import math

import numpy as np
from scipy import special

from .detection import p_one_pmt_acc
from .emission import STD_ANG_DIST_PARS, cherenkov_ang_dist
from .medium import make_cascadia_medium_properties
from .propagation import propagate_distance


def cherenkov_ang_dist_integral(ref_index, lower=-1.0, upper=1.0, pars=STD_ANG_DIST_PARS):
    """∫ cherenkov_ang_dist(x, n) dx over [lower, upper] in closed form
    (the role of _cherenkov_ang_dist_int, src/emission.jl:53-86).  With
    c0 = 1/n and b < 0, ∫_0^U exp(b u^c) du = |b|^(-1/c) γ(1/c, |b| U^c) / c, so
    the integral is a·(J(upper - c0) - J(lower - c0)) + d·(upper - lower) with
    the odd extension J(v) = sign(v)·∫_0^|v|."""
    a, b, c, d = pars
    c0 = 1.0 / np.asarray(ref_index, dtype=np.float64)
    s = 1.0 / c

    def J(v):
        u = np.abs(v)
        return np.sign(v) * abs(b) ** (-s) * special.gamma(s) * special.gammainc(s, abs(b) * u ** c) / c

    return a * (J(upper - c0) - J(lower - c0)) + d * (upper - lower)


# Tabulated like the reference (ChAngDistInt on n = 1.1:0.01:1.5, src/emission.jl:88-99).
_ANG_INT_N = np.round(np.arange(1.1, 1.5 + 1e-9, 0.01), 10)
_ANG_INT_V = cherenkov_ang_dist_integral(_ANG_INT_N)


def cherenkov_ang_dist_int(ref_index):
    """Linear interpolation of the integral over n in [1.1, 1.5]; outside the
    table it raises, as Interpolations' default extrapolation does."""
    n = np.asarray(ref_index, dtype=np.float64)
    if np.any(n < _ANG_INT_N[0] - 1e-12) or np.any(n > _ANG_INT_N[-1] + 1e-12):
        raise ValueError("refractive index outside [1.1, 1.5]")
    return np.interp(n, _ANG_INT_N, _ANG_INT_V)


def get_dir_reweight(thetas, obs_angle, ref_ixs):
    """Weight of each photon for a cascade seen at `obs_angle` from its axis
    (src/modelling.jl:46-49)."""
    thetas = np.asarray(thetas, dtype=np.float64)
    ref_ixs = np.asarray(ref_ixs, dtype=np.float64)
    norm = cherenkov_ang_dist_int(ref_ixs) * 2.0
    return cherenkov_ang_dist(np.cos(thetas - obs_angle), ref_ixs) / norm


def fit_gamma_weighted(x, w, tol=1e-14, max_iter=1000):
    """Weighted maximum-likelihood Gamma(α, θ) (shape, scale), the role of
    Distributions.fit_mle(Gamma, x, w).  Solves log α - ψ(α) = log x̄ - mean(log x)
    by Newton's method in α from Minka's starting point (T. Minka, "Estimating a
    Gamma distribution", 2002); θ = x̄/α."""
    x = np.asarray(x, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    sw = w.sum()
    mx = (w * x).sum() / sw
    mlx = (w * np.log(x)).sum() / sw
    s = math.log(mx) - mlx
    a = (3 - s + math.sqrt((s - 3) ** 2 + 24 * s)) / (12 * s)
    for _ in range(max_iter):
        f = math.log(a) - special.digamma(a) - s
        fp = 1.0 / a - special.polygamma(1, a)
        a_new = a - f / fp
        if a_new <= 0:
            a_new = a / 2
        if abs(a_new - a) <= tol * a:
            a = a_new
            break
        a = a_new
    return a, mx / a


def fit_photon_dist(obs_angles, obs_photon_df, n_ph_gen):
    """Per observation angle: direction re-weighting × absorption weight × PMT
    acceptance, weighted Gamma fit of tres >= 0, detected fraction
    (src/modelling.jl:52-79)."""
    import pandas as pd

    thetas = obs_photon_df["initial_theta"].to_numpy()
    ref_ix = obs_photon_df["ref_ix"].to_numpy()
    abs_w = obs_photon_df["abs_weight"].to_numpy()
    tres = obs_photon_df["tres"].to_numpy()
    pmt = p_one_pmt_acc(obs_photon_df["wavelength"].to_numpy())
    rows = []
    mask = tres >= 0
    for ang in np.ravel(obs_angles):
        total = get_dir_reweight(thetas, ang, ref_ix) * abs_w * pmt
        a, th = fit_gamma_weighted(tres[mask], total[mask])
        rows.append((float(ang), a, th, float(total.sum() / n_ph_gen)))
    return pd.DataFrame(rows, columns=["obs_angle", "fit_alpha", "fit_theta", "det_fraction"])


def sobol_1d(n, lo, hi):
    """The first n points after the origin of the 1-D Sobol sequence (Gray-code
    van der Corput in base 2: 1/2, 3/4, 1/4, 3/8, ...), scaled to [lo, hi] —
    the points SobolSeq([lo], [hi]) yields."""
    from scipy.stats import qmc

    s = qmc.Sobol(d=1, scramble=False)
    m = 1 << int(math.ceil(math.log2(n + 1)))
    pts = s.random(m).ravel()[1:n + 1]
    return lo + (hi - lo) * pts


def make_photon_fits(n_photons_per_dist, n_distances, n_angles, max_dist=300.0, medium=None):
    """Propagate at Sobol-spaced log distances and fit every (distance,
    observation angle) cell (src/modelling.jl:86-111)."""
    import pandas as pd

    medium = medium or make_cascadia_medium_properties()
    distances = 10.0 ** sobol_1d(n_distances, 0.0, math.log10(max_dist))
    obs_angles = sobol_1d(n_angles, 0.0, math.pi)
    out = []
    for dist in distances:
        df = propagate_distance(float(dist), medium, int(n_photons_per_dist))
        if len(df) == 0:
            continue
        fit = fit_photon_dist(obs_angles, df, n_photons_per_dist)
        fit["distance"] = dist
        out.append(fit)
    return pd.concat(out, ignore_index=True) if out else pd.DataFrame()
