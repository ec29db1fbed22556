"""Cascade light-yield generator (src/lightyield.jl): closed forms and
structural properties.  CPU only."""
import math

import numpy as np
import pytest
from scipy import integrate, special

import photonprop as pp
from photonprop import lightyield as ly


def test_dippr105_water_density():
    # ~1000 kg/m³ around 4-25 °C (DDB DIPPR-105 fit for water, src/medium.jl:70)
    assert ly.dippr105(277.15) == pytest.approx(1000.0, rel=0.015)
    assert ly.dippr105(298.15) == pytest.approx(997.0, rel=0.01)
    assert ly.dippr105(298.15) < ly.dippr105(277.15)
    m = pp.make_cascadia_medium_properties()
    assert ly.radiation_length_cm(m) == pytest.approx(36.08 / ly.density(m) * 1000.0)
    assert 30.0 < ly.radiation_length_cm(m) < 40.0


@pytest.mark.parametrize("ptype", list(ly.ParticleType))
@pytest.mark.parametrize("energy", [1.0, 1e3, 1e5])
def test_longitudinal_profile_is_a_density(ptype, energy):
    m = pp.make_cascadia_medium_properties()
    x0 = ly.radiation_length_cm(m)
    tot, _ = integrate.quad(lambda z: ly.longitudinal_profile(energy, z, m, ptype), 0, 200 * x0,
                            limit=400)
    # b·(tb)^(a-1)·e^(-tb)/Γ(a) integrates to 1 over t = z/X0, i.e. to X0 over z
    assert tot == pytest.approx(x0, rel=1e-6)
    # mode at t = (a - 1)/b
    a = ly.long_parameter_a_edep(energy, ptype)
    b = ly.LONG_PARS[ptype].b
    t = np.linspace(0.01, 60, 600001)
    prof = ly.longitudinal_profile(energy, t * x0, m, ptype)
    assert t[np.argmax(prof)] == pytest.approx((a - 1) / b, rel=1e-3)


def test_fractional_contributions_and_track_length():
    m = pp.make_cascadia_medium_properties()
    z = np.linspace(0, 2000, 41)
    f = ly.fractional_contrib_long(1e4, z, m, ly.ParticleType.EMinus)
    assert f[0] == 0 and np.all(f >= 0)
    assert f.sum() == pytest.approx(1.0, rel=1e-10)
    assert ly.cherenkov_track_length(1.0, ly.ParticleType.EMinus) == pytest.approx(532.07078881)
    assert ly.cherenkov_track_length(100.0, ly.ParticleType.Gamma) == pytest.approx(
        532.08540905 * 100 ** 0.99999877)
    assert ly.cherenkov_track_length_dev(4.0, ly.ParticleType.EPlus) == pytest.approx(5.73419669 * 2)


def test_elongated_lightsource_requests():
    m = pp.make_cascadia_medium_properties()
    p = ly.Particle((0.0, 0.0, 0.0), (0.0, 0.0, 2.0), 5.0, 1e3, ly.ParticleType.EMinus)
    tgt = pp.DetectionSphere((5.0, 0.0, 3.0), 0.3)
    reqs = ly.particle_to_elongated_lightsource(p, (0.0, 20.0), 0.5, m, (300.0, 800.0), tgt, seed=4)
    assert reqs.shape == (40,)
    np.testing.assert_allclose(reqs["pos"][:, 2], np.arange(40) * 0.5 + 0.25)
    np.testing.assert_allclose(reqs["t0"], 5.0 + (np.arange(40) * 0.5 + 0.25) / 0.299792458, rtol=1e-6)
    np.testing.assert_allclose(reqs["axis"], [[0, 0, 1.0]] * 40)
    total = (pp.frank_tamm_norm((300.0, 800.0), lambda wl: pp.get_refractive_index(wl, m))
             * ly.cherenkov_track_length(1e3, ly.ParticleType.EMinus))
    assert abs(int(reqs["n_photons"].sum()) - total) < 40  # rounding per source
    assert np.all(reqs["stream"] == np.arange(40))
