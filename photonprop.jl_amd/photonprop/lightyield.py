"""Cascade light yield -> point sources for the batched and serving paths
(mirrors PhotonProp.jl's LightYield module, src/lightyield.jl).

An electromagnetic cascade of energy E (GeV) emits Cherenkov photons along its
axis with the Gamma-shaped longitudinal profile of the radiation depth
(longitudinal_profile, :104-120); its Cherenkov track length is a power law in
E (cherenkov_track_length, :180-183).  particle_to_elongated_lightsource
(:215-285) splits the cascade into point sources on a grid along the axis,
each carrying its share of the total photon count; here the sources come out
as request records ready for Context.propagate_batch or the serving loop.

Units: positions in m, times in ns, the longitudinal grid in m (the reference
mixes cm and m on the same grid); the profile itself is evaluated in cm.

Deviation: the reference's water density DIPPR105 (src/medium.jl:80, :92) is
evaluated with the temperature in °C and with A / B^((1 + (1 - T)/C)^D), which
gives 12.85 kg/m³ for the Cascadia preset and a 28 m radiation length.  Here
the DIPPR-105 correlation rho = A / B^(1 + (1 - T/C)^D) is evaluated in kelvin
(~1011 kg/m³, 0.357 m).
"""
import math
from dataclasses import dataclass
from enum import IntEnum

import numpy as np
from scipy import special

from . import _abi
from .medium import get_refractive_index
from .propagation import C_VACUUM, make_requests
from .spectrum import frank_tamm_norm
from .utils import integrate_gauss_quad


class ParticleType(IntEnum):
    EMinus = 11
    EPlus = -11
    Gamma = 22


@dataclass(frozen=True)
class LongitudinalParameters:
    alpha: float
    beta: float
    b: float


@dataclass(frozen=True)
class CherenkovTrackLengthParameters:
    alpha: float      # cm
    beta: float
    alpha_dev: float  # cm
    beta_dev: float


# Parametrisations (src/lightyield.jl:38-74).
LONG_PARS = {
    ParticleType.EMinus: LongitudinalParameters(2.01849, 1.45469, 0.63207),
    ParticleType.EPlus: LongitudinalParameters(2.00035, 1.45501, 0.63008),
    ParticleType.Gamma: LongitudinalParameters(2.83923, 1.34031, 0.64526),
}
TRACK_PARS = {
    ParticleType.EMinus: CherenkovTrackLengthParameters(532.07078881, 1.00000211, 5.78170887, 0.5),
    ParticleType.EPlus: CherenkovTrackLengthParameters(532.11320598, 0.99999254, 5.73419669, 0.5),
    ParticleType.Gamma: CherenkovTrackLengthParameters(532.08540905, 0.99999877, 5.78170887,
                                                       5.66586567),
}

# DIPPR-105 parameters for water (src/medium.jl:70).
DIPPR105_WATER = (0.14395, 0.0112, 649.727, 0.05107)


def dippr105(temperature_k, params=DIPPR105_WATER):
    """DIPPR-105 liquid density in kg/m³ at `temperature_k` kelvin."""
    a, b, c, d = params
    return a / b ** (1.0 + (1.0 - temperature_k / c) ** d)


def density(medium):
    """Water density in kg/m³ at the medium's temperature (°C)."""
    return dippr105(medium.temperature + 273.15)


def radiation_length_cm(medium):
    return medium.radiation_length / density(medium) * 1000.0


def _pars(p, table):
    return table[ParticleType(p)] if not isinstance(p, (LongitudinalParameters,
                                                       CherenkovTrackLengthParameters)) else p


def long_parameter_a_edep(energy, pars):
    p = _pars(pars, LONG_PARS)
    return p.alpha + p.beta * np.log10(energy)


def longitudinal_profile(energy, z_cm, medium, pars):
    """dE/dz of the cascade per unit depth (z in cm), a Gamma density in the
    radiation depth t = z / X0 (src/lightyield.jl:104-120)."""
    p = _pars(pars, LONG_PARS)
    t = np.asarray(z_cm, dtype=np.float64) / radiation_length_cm(medium)
    a = long_parameter_a_edep(energy, p)
    b = p.b
    return b * ((t * b) ** (a - 1.0) * np.exp(-(t * b)) / special.gamma(a))


def fractional_contrib_long(energy, z_grid_cm, medium, pars):
    """Share of the profile in each grid interval, normalised over the grid's
    range; out[0] = 0 (src/lightyield.jl:123-169)."""
    z = np.asarray(z_grid_cm, dtype=np.float64)
    f = lambda x: longitudinal_profile(energy, x, medium, pars)
    norm = integrate_gauss_quad(f, z[0], z[-1])
    out = np.zeros_like(z)
    for i in range(z.shape[0] - 1):
        out[i + 1] = integrate_gauss_quad(f, z[i], z[i + 1]) / norm
    return out


def cherenkov_track_length(energy, pars):
    """Cherenkov track length in cm (src/lightyield.jl:180-183)."""
    p = _pars(pars, TRACK_PARS)
    return p.alpha * energy ** p.beta


def cherenkov_track_length_dev(energy, pars):
    p = _pars(pars, TRACK_PARS)
    return p.alpha_dev * energy ** p.beta_dev


@dataclass
class Particle:
    """src/lightyield.jl:185-191: position (m), direction, time (ns), energy (GeV)."""

    position: tuple
    direction: tuple
    time: float
    energy: float
    type: ParticleType


def particle_to_elongated_lightsource(particle, len_range_m, precision_m, medium, wl_range,
                                      target, seed=0, max_steps=10, first_stream=0, beta=1.0):
    """Point sources along the cascade axis (src/lightyield.jl:215-285) as
    request records towards `target` (a DetectionSphere): one per grid interval,
    at its midpoint, delayed by the travel time at c, with round(total ·
    fraction) photons on the Cherenkov cone about the cascade direction.
    Returns a REQUEST_DTYPE array."""
    grid_m = np.arange(len_range_m[0], len_range_m[1] + precision_m / 2, precision_m)
    frac = fractional_contrib_long(particle.energy, grid_m * 100.0, medium, particle.type)
    total = (frank_tamm_norm(wl_range, lambda wl: get_refractive_index(wl, medium))
             * cherenkov_track_length(particle.energy, particle.type))
    mid = 0.5 * (grid_m[1:] + grid_m[:-1])
    d = np.asarray(particle.direction, dtype=np.float64)
    d /= np.linalg.norm(d)
    n = mid.shape[0]
    reqs = make_requests(n)
    reqs["pos"] = np.asarray(particle.position, dtype=np.float64)[None, :] + mid[:, None] * d[None, :]
    reqs["t0"] = particle.time + mid / C_VACUUM
    reqs["axis"] = d
    reqs["beta"] = beta
    reqs["n_photons"] = np.round(total * frac[1:]).astype(np.uint64)
    reqs["profile"] = _abi.PP_EMIT_CHERENKOV_CONE
    reqs["center"] = target.position
    reqs["radius"] = target.radius
    reqs["seed"] = seed
    reqs["stream"] = first_stream + np.arange(n, dtype=np.uint32)
    reqs["max_steps"] = max_steps
    return reqs
