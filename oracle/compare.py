"""Photon-by-photon comparison of the GPU library against the oracle.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Both sides draw photon i
from the same Philox stream, so a hit on one side is expected on the other;
they differ only where the device's hardware transcendentals (v_log_f32,
v_sin_f32, v_rcp_f32, ...) round differently from libm, which moves positions
by ~1e-6 relative and can flip a photon that grazes the sphere.
"""
import numpy as np

from . import propagate as oracle_propagate

# Tolerances for hits present on both sides (FP32, ~1 km of path at most).
# The entry distance d = -a - sqrt(a² - (|p-c|² - r²)) cancels for grazing
# rays, so path length (and the weight exp(-L/λ_abs) through it) carries
# ~1e-4 relative FP32 error there; positions stay within 1e-3 m.
TOL = {
    "pos": 2e-3,        # m
    "time": 2e-2,       # ns
    "wavelength": 1e-4,  # nm
    "weight_rel": 5e-4,
    "init_dir": 1e-5,
    "dir": 1e-3,
    "dist_rel": 5e-4,
}


def table_struct(tables):
    """pp_medium_tables for a photonprop.MediumTables; returns (struct, keep-alive)."""
    from photonprop import _abi

    keep = [np.ascontiguousarray(a, dtype=np.float32)
            for a in (tables.n_phase, tables.abs_len, tables.sca_len)]
    t = _abi.pp_medium_tables()
    t.lambda_min, t.lambda_max, t.n_points = tables.lambda_min, tables.lambda_max, tables.n_points
    t.n_phase, t.abs_len, t.sca_len = [a.ctypes.data_as(_abi.c_float_p) for a in keep]
    t.g = tables.g
    return t, keep


def oracle_for(ctx, source, target, seed=0, stream=0, max_steps=10, id_begin=0, id_end=None,
               precision=32, n_threads=None, hits_cap=None):
    """Run the oracle with the same inputs a Context would send to the device.
    `ctx` is a photonprop.Context or a photonprop.MediumTables."""
    from photonprop import _abi
    from photonprop.propagation import emitter_struct, spectrum_struct, sphere_struct

    tables, tkeep = table_struct(getattr(ctx, "tables", ctx))
    spec, keep = spectrum_struct(source.spectrum)
    em = emitter_struct(source)
    tg = sphere_struct(target)
    hits, summ = oracle_propagate(tables, spec, em, tg, seed=seed, stream=stream,
                                  max_steps=max_steps, id_begin=id_begin, id_end=id_end,
                                  precision=precision, n_threads=n_threads, hits_cap=hits_cap,
                                  hit_dtype=_abi.HIT_DTYPE, summary_struct=_abi.pp_summary)
    del keep, tkeep
    return hits, {k: getattr(summ, k) for k, _ in _abi.pp_summary._fields_}


def compare_hits(gpu, cpu):
    """Match hits by photon id; return a dict of mismatch counts and max errors."""
    g = np.sort(gpu, order="photon_id")
    c = np.sort(cpu, order="photon_id")
    common, gi, ci = np.intersect1d(g["photon_id"], c["photon_id"], return_indices=True)
    g, c = g[gi], c[ci]
    out = {
        "n_gpu": int(gpu.shape[0]),
        "n_cpu": int(cpu.shape[0]),
        "n_common": int(common.shape[0]),
        "only_gpu": int(gpu.shape[0] - common.shape[0]),
        "only_cpu": int(cpu.shape[0] - common.shape[0]),
    }
    if common.shape[0] == 0:
        return out
    f = lambda a: a.astype(np.float64)
    out["pos"] = float(np.abs(f(g["pos"]) - f(c["pos"])).max())
    out["time"] = float(np.abs(f(g["time"]) - f(c["time"])).max())
    out["wavelength"] = float(np.abs(f(g["wavelength"]) - f(c["wavelength"])).max())
    out["weight_rel"] = float((np.abs(f(g["weight"]) - f(c["weight"])) /
                               np.maximum(f(c["weight"]), 1e-30)).max())
    out["init_dir"] = float(np.abs(f(g["init_dir"]) - f(c["init_dir"])).max())
    out["dir"] = float(np.abs(f(g["dir"]) - f(c["dir"])).max())
    out["dist_rel"] = float((np.abs(f(g["dist"]) - f(c["dist"])) / np.maximum(f(c["dist"]), 1e-30)).max())
    return out


def assert_parity(cmp, n_hits_ref):
    """Per-photon bar: hit-flag flips at most max(3, 1e-3·hits); every common
    hit within TOL."""
    flips = cmp["only_gpu"] + cmp["only_cpu"]
    assert flips <= max(3, 1e-3 * n_hits_ref), cmp
    for k, tol in TOL.items():
        if k in cmp:
            assert cmp[k] <= tol, (k, cmp[k], tol, cmp)
