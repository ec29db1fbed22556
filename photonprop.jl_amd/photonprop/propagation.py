"""Photon propagation on the GPU through libphotonprop.so.

Mirrors the callable surface of PhotonProp.jl's PhotonPropagationCuda module
(src/photon_prop_cuda.jl): `propagate_distance(distance, medium, n_ph_gen)`
returns the same result-table columns, and `Context.propagate` /
`Context.propagate_batch` expose the general source/target form that the
reference kernel `cuda_propagate_photons!` (:262-386) implements for one
compile-time source.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _abi
from .detection import DetectionSphere
from .emission import AngularEmissionProfile, PhotonSource
from .medium import get_refractive_index, medium_tables
from .spectrum import CherenkovSpectrum, InverseSquareSpectrum, Monochromatic

C_VACUUM = 0.299792458  # m/ns, exact (PhysicalConstants SpeedOfLightInVacuum)
DEFAULT_MAX_STEPS = 10  # src/photon_prop_cuda.jl:319
DEFAULT_HITS_CAP = 1 << 24

REQUEST_DTYPE = np.dtype({
    "names": ["pos", "t0", "axis", "beta", "n_photons", "profile", "reserved",
              "center", "radius", "seed", "stream", "max_steps"],
    "formats": [("<f4", (3,)), "<f4", ("<f4", (3,)), "<f4", "<u8", "<u4", "<u4",
                ("<f4", (3,)), "<f4", "<u8", "<u4", "<u4"],
    "offsets": [0, 12, 16, 28, 32, 40, 44, 48, 60, 64, 72, 76],
    "itemsize": 80,
})
assert REQUEST_DTYPE.itemsize == ctypes.sizeof(_abi.pp_request)


def spectrum_struct(spectrum):
    """Build a pp_spectrum; returns (struct, keep-alive array or None).
    Unknown spectrum types raise ValueError like the reference's dispatch
    fallback (src/photon_prop_cuda.jl:70)."""
    s = _abi.pp_spectrum()
    keep = None
    if isinstance(spectrum, Monochromatic):
        s.kind = _abi.PP_SPEC_MONO
        s.mono = float(spectrum.wavelength)
    elif isinstance(spectrum, CherenkovSpectrum):
        keep = np.ascontiguousarray(spectrum.knots, dtype=np.float32)
        s.kind = _abi.PP_SPEC_KNOTS
        s.n_knots = keep.shape[0]
        s.knots = keep.ctypes.data_as(_abi.c_float_p)
    elif isinstance(spectrum, InverseSquareSpectrum):
        s.kind = _abi.PP_SPEC_INV_LAMBDA2
        s.lo = float(spectrum.lo)
        s.hi = float(spectrum.hi)
    else:
        raise ValueError("Cannot initialize %s" % type(spectrum).__name__)
    return s, keep


def emitter_struct(source):
    e = _abi.pp_emitter()
    e.pos[:] = [float(v) for v in source.position]
    e.axis[:] = [float(v) for v in source.direction]
    e.t0 = float(source.time)
    e.beta = float(source.beta)
    e.n_photons = int(source.photons)
    e.profile = source.emission_profile.code
    return e


def sphere_struct(target):
    t = _abi.pp_sphere()
    t.center[:] = [float(v) for v in target.position]
    t.radius = float(target.radius)
    return t


def summary_dict(s):
    return {k: getattr(s, k) for k, _ in _abi.pp_summary._fields_ if k != "reserved"}


@dataclass
class PropagationResult:
    """Hits (numpy records of include/photonprop.h pp_hit) and the FP64 summary."""

    hits: np.ndarray
    summary: dict


class Context:
    """One device context: medium tables resident on the GPU, one HIP stream
    (pp_create).  Not safe for concurrent use from several threads."""

    def __init__(self, medium, device=0, lambda_range=None, n_points=None):
        self.lib = _abi.load_library()
        self.medium = medium
        self.tables = medium_tables(medium, lambda_range, n_points)
        self._handle = ctypes.c_void_p()
        t = self._table_struct()
        _abi.check(self.lib.pp_create(ctypes.byref(self._handle), int(device), ctypes.byref(t)),
                   self.lib)

    def _table_struct(self):
        tb = self.tables
        self._tkeep = [np.ascontiguousarray(a, dtype=np.float32)
                       for a in (tb.n_phase, tb.abs_len, tb.sca_len)]
        t = _abi.pp_medium_tables()
        t.lambda_min = tb.lambda_min
        t.lambda_max = tb.lambda_max
        t.n_points = tb.n_points
        t.n_phase, t.abs_len, t.sca_len = [a.ctypes.data_as(_abi.c_float_p) for a in self._tkeep]
        t.g = tb.g
        return t

    @property
    def handle(self):
        if not self._handle:
            raise RuntimeError("context destroyed")
        return self._handle

    def set_medium(self, medium, lambda_range=None, n_points=None):
        self.medium = medium
        self.tables = medium_tables(medium, lambda_range, n_points)
        t = self._table_struct()
        _abi.check(self.lib.pp_set_medium(self.handle, ctypes.byref(t)), self.lib)

    def close(self):
        if self._handle:
            self.lib.pp_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self):
        name = ctypes.create_string_buffer(256)
        ncu = ctypes.c_int()
        _abi.check(self.lib.pp_device_info(self.handle, name, 256, ctypes.byref(ncu)), self.lib)
        return name.value.decode(), ncu.value

    def propagate(self, source, target, seed=0, stream=0, max_steps=DEFAULT_MAX_STEPS,
                  hits_cap=None, allow_overflow=False):
        """Propagate source.photons photons towards one target (pp_propagate).
        Raises PhotonPropError on overflow unless allow_overflow."""
        cap = int(min(source.photons, DEFAULT_HITS_CAP) if hits_cap is None else hits_cap)
        hits = np.empty(cap, dtype=_abi.HIT_DTYPE)
        spec, keep = spectrum_struct(source.spectrum)
        em = emitter_struct(source)
        tg = sphere_struct(target)
        summ = _abi.pp_summary()
        rc = self.lib.pp_propagate(self.handle, ctypes.byref(em), ctypes.byref(spec),
                                   ctypes.byref(tg), int(seed), int(stream), int(max_steps),
                                   _abi.ptr(hits), cap, ctypes.byref(summ))
        del keep
        if not (rc == _abi.PP_ERR_OVERFLOW and allow_overflow):
            _abi.check(rc, self.lib)
        return PropagationResult(hits[: summ.n_stored], summary_dict(summ))

    def propagate_batch(self, requests, spectrum, hits_cap=None, allow_overflow=False,
                        per_request_cap=0):
        """Propagate many requests in one persistent launch (pp_propagate_batch).
        `requests` is a REQUEST_DTYPE array.  Returns (hits, offsets, summaries):
        request r owns hits[offsets[r]:offsets[r+1]].  per_request_cap > 0 keeps
        at most that many records per request.  A request's records come in the
        order the kernel's waves flushed them, which varies from run to run, and
        so does the subset a cap keeps; the summaries are the same on every run."""
        reqs = np.ascontiguousarray(requests, dtype=REQUEST_DTYPE)
        n = reqs.shape[0]
        cap = int(min(int(reqs["n_photons"].sum()), DEFAULT_HITS_CAP)
                  if hits_cap is None else hits_cap)
        out_cap = min(cap, per_request_cap * n) if per_request_cap else cap
        hits = np.empty(out_cap, dtype=_abi.HIT_DTYPE)
        offsets = np.zeros(n + 1, dtype=np.uint64)
        sums = np.zeros(n, dtype=_abi.SUMMARY_DTYPE)
        spec, keep = spectrum_struct(spectrum)
        rc = self.lib.pp_propagate_batch(
            self.handle, reqs.ctypes.data_as(ctypes.POINTER(_abi.pp_request)), n,
            ctypes.byref(spec), _abi.ptr(hits) if out_cap else None, cap, int(per_request_cap),
            _abi.ptr(offsets), _abi.ptr(sums))
        del keep
        if not (rc == _abi.PP_ERR_OVERFLOW and allow_overflow):
            _abi.check(rc, self.lib)
        return hits[: int(offsets[-1])], offsets, sums

    def last_kernel_ms(self):
        ms = ctypes.c_float()
        _abi.check(self.lib.pp_last_kernel_ms(self.handle, ctypes.byref(ms)), self.lib)
        return ms.value


def make_requests(n):
    """Zeroed request array of length n with max_steps = 10."""
    r = np.zeros(n, dtype=REQUEST_DTYPE)
    r["max_steps"] = DEFAULT_MAX_STEPS
    return r


def request_from(source, target, seed=0, stream=0, max_steps=DEFAULT_MAX_STEPS):
    r = make_requests(1)
    r["pos"][0] = source.position
    r["axis"][0] = source.direction
    r["t0"][0] = source.time
    r["beta"][0] = source.beta
    r["n_photons"][0] = source.photons
    r["profile"][0] = source.emission_profile.code
    r["center"][0] = target.position
    r["radius"][0] = target.radius
    r["seed"][0] = seed
    r["stream"][0] = stream
    r["max_steps"][0] = max_steps
    return r


_default_ctx = {}


def _ctx_for(medium, device=0):
    key = (id(medium), device)
    ctx = _default_ctx.get(key)
    if ctx is None or ctx.medium is not medium:
        ctx = Context(medium, device=device)
        _default_ctx[key] = ctx
    return ctx


def result_table(hits, medium, distance, target_radius):
    """The reference's post-processing columns (src/photon_prop_cuda.jl:622-643):
    tres with the geometric time at n(800 nm), initial θ, ref_ix, abs_weight,
    dist_travelled, wavelength."""
    import pandas as pd

    wl = hits["wavelength"].astype(np.float64)
    dist = hits["dist"].astype(np.float64)
    ref_ix = np.asarray(get_refractive_index(wl, medium), dtype=np.float64)
    photon_times = dist / (C_VACUUM / ref_ix)
    n800 = float(get_refractive_index(800.0, medium))
    tgeo = (distance - target_radius) / (C_VACUUM / n800)
    return pd.DataFrame({
        "tres": photon_times - tgeo,
        "initial_theta": np.arccos(np.clip(hits["init_dir"][:, 2].astype(np.float64), -1, 1)),
        "ref_ix": ref_ix,
        "abs_weight": hits["weight"].astype(np.float64),
        "dist_travelled": dist,
        "wavelength": wl,
    })


def propagate_distance(distance, medium, n_ph_gen, seed=0, device=0):
    """Drop-in for propagate_distance (src/photon_prop_cuda.jl:588-644): a
    point source at the origin with the reference's CherenkovSpectrum((300, 800),
    20) and :IsotropicEmission, a sphere of radius 0.21 m at (0, 0, distance).
    Returns an empty table when no photon arrives (:611-614)."""
    import pandas as pd

    target_radius = 0.21
    spectrum = CherenkovSpectrum((300.0, 800.0), 20, medium)
    source = PhotonSource((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0.0, int(n_ph_gen), spectrum,
                          AngularEmissionProfile("IsotropicEmission"))
    target = DetectionSphere((0.0, 0.0, float(distance)), target_radius)
    ctx = _ctx_for(medium, device)
    res = ctx.propagate(source, target, seed=seed)
    if res.hits.shape[0] == 0:
        return pd.DataFrame()
    return result_table(res.hits, medium, float(distance), target_radius)
