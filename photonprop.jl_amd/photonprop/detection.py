"""Targets and PMT acceptance (mirrors PhotonProp.jl's Detection module,
src/detection.jl)."""
import os
from dataclasses import dataclass

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "pmt_acceptance.csv")


class PhotonTarget:
    """Abstract target (src/detection.jl:10)."""


@dataclass
class DetectionSphere(PhotonTarget):
    """Spherical target: centre (m) and radius (m) (src/detection.jl:12-15)."""

    position: tuple
    radius: float


class PMTWavelengthAcceptance:
    """Linear interpolation of a PMT's acceptance versus wavelength in nm
    (src/detection.jl:18-28); outside the tabulated range it raises, as the
    reference's LinearInterpolation does."""

    def __init__(self, xs, ys):
        self.xs = np.asarray(xs, dtype=np.float64)
        self.ys = np.asarray(ys, dtype=np.float64)
        if self.xs.ndim != 1 or self.xs.shape != self.ys.shape or np.any(np.diff(self.xs) <= 0):
            raise ValueError("acceptance table must be 1-D, matching and strictly increasing")

    def __call__(self, wavelength):
        w = np.asarray(wavelength, dtype=np.float64)
        if np.any(w < self.xs[0]) or np.any(w > self.xs[-1]):
            raise ValueError("wavelength outside [%g, %g] nm" % (self.xs[0], self.xs[-1]))
        return np.interp(w, self.xs, self.ys)


def load_pmt_acceptance(path=_DATA):
    """Two-column CSV (wavelength nm, acceptance) without header, as the
    reference's PMTAcc.csv (read at src/detection.jl:31)."""
    d = np.loadtxt(path, delimiter=",", dtype=np.float64)
    return PMTWavelengthAcceptance(d[:, 0], d[:, 1])


p_one_pmt_acc = load_pmt_acceptance()
