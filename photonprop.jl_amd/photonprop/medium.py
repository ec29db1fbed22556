"""Media: the reference's water model and the tabulated form the device uses.

Mirrors PhotonProp.jl's Medium module (src/medium.jl).  The reference passes a
medium into its kernel as a compile-time constant (Val(medium),
src/photon_prop_cuda.jl:274, :609); here every medium is turned into tables
of n, λ_abs and λ_sca on a uniform wavelength grid plus the Henyey-Greenstein
g, uploaded once per context.
"""
from dataclasses import dataclass

import numpy as np

from .utils import fast_linear_interp

BAR_PER_ATM = 1.01325


class MediumProperties:
    """Abstract medium (src/medium.jl:15).  Subclasses provide
    get_refractive_index / get_scattering_length / get_absorption_length over
    wavelengths in nm, the HG asymmetry `hg_g` and a default table range."""

    hg_g = 0.0
    table_range = (300.0, 600.0)
    table_points = 1025

    def refractive_index(self, wavelength):
        raise NotImplementedError("Not implemented for %s" % type(self).__name__)

    def scattering_length(self, wavelength):
        raise NotImplementedError("Not implemented for %s" % type(self).__name__)

    def absorption_length(self, wavelength):
        raise NotImplementedError("Not implemented for %s" % type(self).__name__)


@dataclass
class WaterProperties(MediumProperties):
    """Water medium (src/medium.jl:23-50).  Units: salinity permille, pressure
    atm, temperature °C, particle concentrations ppm, radiation length g/cm².

    `hg_g` is not a field of the reference struct: the reference kernel
    hard-codes g = 0.99 for every medium (src/photon_prop_cuda.jl:229), so that
    is the value used for this medium."""

    salinity: float
    pressure: float
    temperature: float
    vol_conc_small_part: float
    vol_conc_large_part: float
    radiation_length: float
    hg_g: float = 0.99
    table_range: tuple = (300.0, 800.0)
    table_points: int = 1001

    def refractive_index(self, wavelength):
        return get_refractive_index(wavelength, self)

    def scattering_length(self, wavelength):
        return get_scattering_length(wavelength, self)

    def absorption_length(self, wavelength):
        return get_absorption_length(wavelength, self)


def make_cascadia_medium_properties(T=float):
    """Cascadia Basin preset (src/medium.jl:52-58); pressure 269.44088 bar."""
    return WaterProperties(
        salinity=34.82, pressure=269.44088 / BAR_PER_ATM, temperature=1.8,
        vol_conc_small_part=0.0075, vol_conc_large_part=0.0075, radiation_length=36.08)


def salinity(m):
    return m.salinity


def pressure(m):
    return m.pressure


def temperature(m):
    return m.temperature


def vol_conc_small_part(m):
    return m.vol_conc_small_part


def vol_conc_large_part(m):
    return m.vol_conc_large_part


def radiation_length(m):
    return m.radiation_length


def get_refractive_index_fry(wavelength, salinity, temperature, pressure):
    """Quan & Fry (Appl. Opt. 34 (1995) 3477) phase index of sea water with the
    pressure term of Schuster (2002) (src/medium.jl:125-158).  wavelength nm,
    salinity permille, temperature °C, pressure atm."""
    n0, n1, n2, n3, n4 = 1.31405, 1.45e-5, 1.779e-4, 1.05e-6, 1.6e-8
    n5, n6, n7, n8, n9, n10 = 2.02e-6, 15.868, 0.01155, 0.00423, 4382.0, 1.1455e6
    a01 = (n0 + (n2 - n3 * temperature + n4 * temperature ** 2) * salinity
           - n5 * temperature ** 2 + n1 * pressure)
    a2 = n6 + n7 * salinity - n8 * temperature
    a3 = -n9
    a4 = n10
    x = 1.0 / np.asarray(wavelength, dtype=np.float64)
    return a01 + x * (a2 + x * (a3 + x * a4))


def get_refractive_index(wavelength, medium):
    """src/medium.jl:179-183."""
    if isinstance(medium, WaterProperties):
        return get_refractive_index_fry(wavelength, medium.salinity, medium.temperature,
                                        medium.pressure)
    return medium.refractive_index(wavelength)


def get_sca_len_part_conc(wavelength, vol_conc_small_part, vol_conc_large_part):
    """Scattering length in m from small/large particle concentrations (ppm),
    clsim parametrisation (src/medium.jl:201-214)."""
    x = 550.0 / np.asarray(wavelength, dtype=np.float64)
    coeff = (0.0017 * x ** 4.3 + 1.34 * vol_conc_small_part * x ** 1.7
             + 0.312 * vol_conc_large_part * x ** 0.3)
    return 1.0 / coeff


def get_scattering_length(wavelength, medium):
    """src/medium.jl:233-235."""
    if isinstance(medium, WaterProperties):
        return get_sca_len_part_conc(wavelength, medium.vol_conc_small_part,
                                     medium.vol_conc_large_part)
    return medium.scattering_length(wavelength)


ABS_KNOTS_NM = np.array([300.0, 365.0, 400.0, 450.0, 585.0, 800.0])
ABS_KNOTS_M = np.array([10.4, 10.4, 14.5, 27.7, 7.1, 7.1])


def get_absorption_length(wavelength, medium):
    """Six-knot piecewise-linear absorption length, clamped (src/medium.jl:243-251)."""
    if isinstance(medium, WaterProperties):
        return fast_linear_interp(wavelength, ABS_KNOTS_NM, ABS_KNOTS_M)
    return medium.absorption_length(wavelength)


@dataclass
class SeawaterConfigMedium(MediumProperties):
    """Seawater-like medium with closed-form optics:
      n(λ)     linear from n_300 at 300 nm to n_600 at 600 nm,
      λ_abs(λ) abs_peak m at 450 nm falling linearly to abs_edge m at 300 and 600 nm,
      λ_sca(λ) sca_450·(λ/450 nm)^4 m,
      HG g = hg_g.
    Defaults: 1.36 → 1.34, 30 m / 10 m, 50 m, g = 0.92."""

    n_300: float = 1.36
    n_600: float = 1.34
    abs_peak: float = 30.0
    abs_edge: float = 10.0
    sca_450: float = 50.0
    hg_g: float = 0.92
    table_range: tuple = (300.0, 600.0)
    table_points: int = 1025

    def refractive_index(self, wavelength):
        w = np.asarray(wavelength, dtype=np.float64)
        return self.n_300 + (self.n_600 - self.n_300) * (w - 300.0) / 300.0

    def absorption_length(self, wavelength):
        w = np.clip(np.asarray(wavelength, dtype=np.float64), 300.0, 600.0)
        return self.abs_peak - (self.abs_peak - self.abs_edge) * np.abs(w - 450.0) / 150.0

    def scattering_length(self, wavelength):
        w = np.asarray(wavelength, dtype=np.float64)
        return self.sca_450 * (w / 450.0) ** 4


@dataclass
class MediumTables:
    """Uniform-grid tables as uploaded to the device (pp_medium_tables)."""

    lambda_min: float
    lambda_max: float
    n_phase: np.ndarray
    abs_len: np.ndarray
    sca_len: np.ndarray
    g: float

    @property
    def n_points(self):
        return int(self.n_phase.shape[0])

    @property
    def grid(self):
        return np.linspace(self.lambda_min, self.lambda_max, self.n_points)


def medium_tables(medium, lambda_range=None, n_points=None):
    """Tabulate a medium on a uniform grid (FP64 evaluation, FP32 tables)."""
    if isinstance(medium, MediumTables):
        return medium
    lo, hi = lambda_range or medium.table_range
    n = int(n_points or medium.table_points)
    grid = np.linspace(lo, hi, n)
    f32 = lambda a: np.ascontiguousarray(np.broadcast_to(a, grid.shape), dtype=np.float32)
    return MediumTables(
        lambda_min=float(lo), lambda_max=float(hi),
        n_phase=f32(get_refractive_index(grid, medium)),
        abs_len=f32(get_absorption_length(grid, medium)),
        sca_len=f32(get_scattering_length(grid, medium)),
        g=float(medium.hg_g))
