"""Targets (mirrors PhotonProp.jl's Detection module, src/detection.jl)."""
from dataclasses import dataclass


class PhotonTarget:
    """Abstract target (src/detection.jl:10)."""


@dataclass
class DetectionSphere(PhotonTarget):
    """Spherical target: centre (m) and radius (m) (src/detection.jl:12-15)."""

    position: tuple
    radius: float
