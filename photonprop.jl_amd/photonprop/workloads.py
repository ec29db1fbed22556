"""Synthetic request sets for the scale, batched and serving workloads.

The reference has no batching or serving layer (its only batching is the
unfinished split_source / propagate_sources, src/photon_prop_cuda.jl:474-531);
these generators define the workloads the library is measured on.
"""
import math

import numpy as np

from . import _abi
from .propagation import make_requests


def _unit_sphere(rng, n):
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2 * math.pi, n)
    s = np.sqrt(1.0 - z * z)
    return np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)


def scale_geometries(photons=1_000_000_000, radius=0.21, seed=0,
                     distances=None, angles_deg=(0.0, 45.0, 90.0, 135.0)):
    """16 distances geomspace(1, 100, 16) m x directions at 0/45/90/135 degrees
    from the +z emitter axis; one request (stream = index) per geometry."""
    distances = np.geomspace(1.0, 100.0, 16) if distances is None else np.asarray(distances)
    reqs = make_requests(len(distances) * len(angles_deg))
    i = 0
    for d in distances:
        for a in angles_deg:
            ar = math.radians(a)
            reqs["axis"][i] = (0.0, 0.0, 1.0)
            reqs["beta"][i] = 1.0
            reqs["n_photons"][i] = photons
            reqs["profile"][i] = _abi.PP_EMIT_CHERENKOV_CONE
            reqs["center"][i] = (d * math.sin(ar), 0.0, d * math.cos(ar))
            reqs["radius"][i] = radius
            reqs["seed"][i] = seed
            reqs["stream"][i] = i
            i += 1
    return reqs


def batched_pairs(n_pairs=4096, seed=0, photons=(1e3, 1e7), distance=(1.0, 200.0),
                  radius=(0.2, 0.5)):
    """Uneven batch: photon budgets log-uniform on `photons`, distances
    log-uniform on `distance` m, emitter axes uniform on the sphere, radii
    uniform on `radius` m.  The emitter sits at the origin and every target
    centre lies on +x at its distance (the axis is what varies).  Request r
    uses Philox stream r of key `seed`."""
    rng = np.random.default_rng(seed)
    reqs = make_requests(n_pairs)
    reqs["n_photons"] = np.round(10.0 ** rng.uniform(math.log10(photons[0]),
                                                     math.log10(photons[1]), n_pairs)).astype(np.uint64)
    d = 10.0 ** rng.uniform(math.log10(distance[0]), math.log10(distance[1]), n_pairs)
    reqs["center"][:, 0] = d
    reqs["axis"] = _unit_sphere(rng, n_pairs)
    reqs["radius"] = rng.uniform(radius[0], radius[1], n_pairs)
    reqs["beta"] = 1.0
    reqs["profile"] = _abi.PP_EMIT_CHERENKOV_CONE
    reqs["seed"] = seed
    reqs["stream"] = np.arange(n_pairs, dtype=np.uint32)
    return reqs


def serving_requests(n, rate_hz, seed=0, photons=(1e4, 1e8), distance=(1.0, 150.0),
                     radius=0.21):
    """Poisson arrivals at rate_hz: returns (arrival times in s, requests).
    Budgets log-uniform on `photons`, random emitter axis, target at a
    log-uniform distance along a random direction."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_hz, n))
    reqs = make_requests(n)
    reqs["n_photons"] = np.round(10.0 ** rng.uniform(math.log10(photons[0]),
                                                     math.log10(photons[1]), n)).astype(np.uint64)
    d = 10.0 ** rng.uniform(math.log10(distance[0]), math.log10(distance[1]), n)
    reqs["center"] = _unit_sphere(rng, n) * d[:, None]
    reqs["axis"] = _unit_sphere(rng, n)
    reqs["radius"] = radius
    reqs["beta"] = 1.0
    reqs["profile"] = _abi.PP_EMIT_CHERENKOV_CONE
    reqs["seed"] = seed + 1
    reqs["stream"] = np.arange(n, dtype=np.uint32)
    return t, reqs
