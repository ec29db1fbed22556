"""Wavelength spectra (mirrors PhotonProp.jl's Spectral module, src/spectrum.jl)."""
from dataclasses import dataclass, field

import numpy as np

from .medium import get_refractive_index
from .utils import integrate_gauss_quad

# CODATA 2018 fine-structure constant (PhysicalConstants.CODATA2018, src/spectrum.jl:33).
FINE_STRUCTURE_CONSTANT = 7.2973525693e-3


class Spectrum:
    """Abstract spectrum (src/spectrum.jl:19)."""


@dataclass
class Monochromatic(Spectrum):
    """src/spectrum.jl:21-23."""

    wavelength: float


@dataclass
class InverseSquareSpectrum(Spectrum):
    """dN/dλ ∝ 1/λ² on [lo, hi] nm, sampled by its closed-form inverse CDF
    1/λ = 1/lo - u·(1/lo - 1/hi).  This is the Frank-Tamm shape with a
    wavelength-independent refractive index."""

    lo: float = 300.0
    hi: float = 600.0

    def mean(self):
        return self.lo * self.hi * np.log(self.hi / self.lo) / (self.hi - self.lo)

    def cdf(self, wl):
        wl = np.clip(np.asarray(wl, dtype=np.float64), self.lo, self.hi)
        return (1.0 / self.lo - 1.0 / wl) / (1.0 / self.lo - 1.0 / self.hi)


def frank_tamm(wavelength, ref_index):
    """Frank-Tamm photon density per unit length and wavelength, nm⁻²
    (src/spectrum.jl:32-34)."""
    wavelength = np.asarray(wavelength, dtype=np.float64)
    return 2 * np.pi * FINE_STRUCTURE_CONSTANT / wavelength ** 2 * (1 - 1 / np.asarray(ref_index) ** 2)


def frank_tamm_norm(wl_range, ref_index_func):
    """Cherenkov photons per cm in wl_range (src/spectrum.jl:42-45)."""
    f = lambda x: frank_tamm(x, ref_index_func(x))
    return integrate_gauss_quad(f, wl_range[0], wl_range[1]) * 1e7


def frank_tamm_inverted_cdf(wl_range, medium, step_size=1.0):
    """CDF of the Frank-Tamm spectrum on a step_size grid, returned as
    (sorted cdf values, wavelengths) (src/spectrum.jl:54-70)."""
    lo, hi = float(wl_range[0]), float(wl_range[1])
    wl_steps = np.arange(lo, hi + step_size / 2, step_size)
    n_func = lambda wl: get_refractive_index(wl, medium)
    full = frank_tamm_norm((lo, hi), n_func)
    norms = np.zeros(wl_steps.shape[0])
    for i in range(1, wl_steps.shape[0]):
        norms[i] = frank_tamm_norm((lo, wl_steps[i]), n_func) / full
    order = np.argsort(norms, kind="stable")
    return norms[order], wl_steps[order]


@dataclass
class CherenkovSpectrum(Spectrum):
    """Inverse-CDF knots equidistant in u on [0, 1] (src/spectrum.jl:90-103):
    CherenkovSpectrum((300, 800), 20, medium)."""

    wl_range: tuple
    interp_steps: int
    medium: object = None
    knots: np.ndarray = field(default=None, repr=False)

    def __post_init__(self):
        if self.knots is None:
            norms, wls = frank_tamm_inverted_cdf(self.wl_range, self.medium)
            u = np.linspace(0.0, 1.0, int(self.interp_steps))
            self.knots = np.interp(u, norms, wls)
        self.knots = np.asarray(self.knots, dtype=np.float64)
