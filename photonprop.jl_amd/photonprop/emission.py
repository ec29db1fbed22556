"""Emitters (mirrors PhotonProp.jl's Emission module, src/emission.jl)."""
from dataclasses import dataclass

import numpy as np

from . import _abi

# Profile tags.  "IsotropicEmission" is the reference's tag and behaves as the
# reference kernel does (θ ~ U(-1, 1) radians, φ ~ U(0, 2π), emitter axis
# ignored; src/photon_prop_cuda.jl:52-56).  "CherenkovCone" emits on the cone
# cosθ = 1/(n(λ)·β) about the emitter axis; "UniformSphere" is isotropic.
_PROFILE_CODES = {
    "CherenkovCone": _abi.PP_EMIT_CHERENKOV_CONE,
    "UniformSphere": _abi.PP_EMIT_ISOTROPIC,
    "IsotropicEmission": _abi.PP_EMIT_REF_ISOTROPIC,
}


@dataclass(frozen=True)
class AngularEmissionProfile:
    """Emission-profile tag (src/emission.jl:13, AngularEmissionProfile{U,T})."""

    kind: str = "CherenkovCone"

    @property
    def code(self):
        try:
            return _PROFILE_CODES[self.kind]
        except KeyError:
            # The reference's dispatch fallback (src/photon_prop_cuda.jl:67).
            raise ValueError("Cannot initialize %s" % self.kind) from None


@dataclass
class PhotonSource:
    """Point emitter (src/emission.jl:102-109).  `beta` (particle speed / c,
    default 1) sets the Cherenkov cone; the reference struct has no such field."""

    position: tuple
    direction: tuple
    time: float
    photons: int
    spectrum: object
    emission_profile: AngularEmissionProfile = AngularEmissionProfile()
    beta: float = 1.0


# Cherenkov angular distribution of EM cascades (arXiv:1210.5140), parameters
# for electrons (src/emission.jl:15-23).
STD_ANG_DIST_PARS = (4.27033, -6.02527, 0.29887, -0.00103)


def cherenkov_ang_dist(costheta, ref_index, pars=STD_ANG_DIST_PARS):
    """a·exp(b·|cosθ - 1/n|^c) + d (src/emission.jl:32-44)."""
    a, b, c, d = pars
    costheta = np.asarray(costheta, dtype=np.float64)
    return a * np.exp(b * np.abs(costheta - 1.0 / ref_index) ** c) + d
