"""CPU-side checks of the C-ABI library and the host mirror (no GPU compute)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import photonprop as pp
from photonprop import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "photonprop.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load_library()
    names = declared_functions()
    assert "pp_propagate" in names and "pp_create" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_abi.SIGNATURES), set(names) ^ set(_abi.SIGNATURES)
    assert lib.pp_abi_version() == 1


def test_julia_binding_calls_only_declared_symbols():
    """photonprop.jl_amd/julia/PhotonPropHIP.jl (not executable here: no Julia)
    ccalls only functions that the header declares and the library exports."""
    jl = open(os.path.join(ROOT, "photonprop.jl_amd", "julia", "PhotonPropHIP.jl")).read()
    used = set(re.findall(r"\(:(pp_[a-z0-9_]+), LIB\)", jl))
    assert {"pp_create", "pp_propagate", "pp_propagate_batch", "pp_server_submit"} <= used
    lib = _abi.load_library()
    declared = set(declared_functions())
    for n in sorted(used):
        assert n in declared and hasattr(lib, n), n


def test_loadgen_library_loads_and_rejects_bad_arguments():
    """The serving load generator (libpploadgen.so) loads beside the library,
    exports its driver, and refuses a NULL server without calling anything."""
    from photonprop import serving

    lg = serving._loadgen()
    assert hasattr(lg, "ppl_open_loop") and hasattr(lg, "ppl_per_request_loop")
    mk = ctypes.c_double()
    rc = lg.ppl_open_loop(None, None, None, None, None, None, 0, 0, 1.0, None, None, None,
                          ctypes.byref(mk))
    assert rc == serving.LOADGEN_BAD_ARG
    rc = lg.ppl_per_request_loop(None, None, None, None, None, 0, 1.0, None, None,
                                 ctypes.byref(mk))
    assert rc == serving.LOADGEN_BAD_ARG


def test_struct_layouts_match_header():
    # sizes fixed by the header's comments / layout
    assert ctypes.sizeof(_abi.pp_summary) == 64
    assert _abi.HIT_DTYPE.itemsize == 64
    assert ctypes.sizeof(_abi.pp_emitter) == 48
    assert ctypes.sizeof(_abi.pp_sphere) == 16
    assert ctypes.sizeof(_abi.pp_request) == 80
    assert ctypes.sizeof(_abi.pp_spectrum) == 32
    assert ctypes.sizeof(_abi.pp_medium_tables) == 48


def test_header_compiles_as_c_and_sizes_agree(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include "photonprop.h"\n#include <stdio.h>\nint main(void){printf("%zu %zu %zu %zu %zu %zu %zu\\n",'
                   'sizeof(pp_hit),sizeof(pp_summary),sizeof(pp_emitter),sizeof(pp_sphere),'
                   'sizeof(pp_request),sizeof(pp_spectrum),sizeof(pp_medium_tables));return 0;}\n')
    exe = tmp_path / "sz"
    import subprocess
    subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                           str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)]).decode().split()
    assert [int(x) for x in out] == [64, 64, 48, 16, 80, 32, 48]


def test_errors_surface_without_gpu():
    lib = _abi.load_library()
    # NULL output pointer is rejected before any device call
    t = _abi.pp_medium_tables()
    rc = lib.pp_create(None, 0, ctypes.byref(t))
    assert rc == _abi.PP_ERR_INVALID_ARG
    assert "NULL" in lib.pp_last_error().decode()
    rc = lib.pp_propagate(None, None, None, None, 0, 0, 10, None, 0, None)
    assert rc == _abi.PP_ERR_INVALID_ARG
    with pytest.raises(pp.PhotonPropError):
        _abi.check(rc, lib)


def test_unsupported_tags_raise_like_reference():
    with pytest.raises(ValueError, match="Cannot initialize"):
        pp.AngularEmissionProfile("Lambertian").code
    from photonprop.propagation import spectrum_struct
    with pytest.raises(ValueError, match="Cannot initialize"):
        spectrum_struct(object())


# ------------------------------------------------------------ host mirror ----

def test_config_medium_closed_forms():
    m = pp.SeawaterConfigMedium()
    assert m.refractive_index(300.0) == pytest.approx(1.36)
    assert m.refractive_index(600.0) == pytest.approx(1.34)
    assert m.absorption_length(450.0) == pytest.approx(30.0)
    assert m.absorption_length(300.0) == pytest.approx(10.0)
    assert m.absorption_length(600.0) == pytest.approx(10.0)
    assert m.scattering_length(450.0) == pytest.approx(50.0)
    assert m.scattering_length(300.0) == pytest.approx(50 * (300 / 450) ** 4)
    t = pp.medium_tables(m)
    assert t.n_points == 1025 and t.grid[512] == 450.0  # the absorption kink is a grid point
    assert t.g == pytest.approx(0.92)


def test_water_properties_reference_values():
    m = pp.make_cascadia_medium_properties()
    # Quan-Fry at 400 nm for the cascadia preset, evaluated by hand from the
    # polynomial's coefficients with S=34.82, T=1.8, p=265.92 atm.
    S, T, p = 34.82, 1.8, 269.44088 / 1.01325
    a01 = 1.31405 + (1.779e-4 - 1.05e-6 * T + 1.6e-8 * T * T) * S - 2.02e-6 * T * T + 1.45e-5 * p
    a2 = 15.868 + 0.01155 * S - 0.00423 * T
    x = 1 / 400.0
    assert pp.get_refractive_index(400.0, m) == pytest.approx(a01 + a2 * x - 4382 * x * x + 1.1455e6 * x ** 3)
    # refractive index of sea water is ~1.35 in the blue and falls with λ
    n = pp.get_refractive_index(np.array([350.0, 450.0, 550.0]), m)
    assert 1.33 < n[-1] < n[1] < n[0] < 1.37
    # absorption knots (src/medium.jl:245-246) and clamping
    assert pp.get_absorption_length(450.0, m) == pytest.approx(27.7)
    assert pp.get_absorption_length(200.0, m) == pytest.approx(10.4)
    assert pp.get_absorption_length(1000.0, m) == pytest.approx(7.1)
    assert pp.get_absorption_length(500.0, m) == pytest.approx(27.7 + (7.1 - 27.7) * 50 / 135)
    # pure-water Rayleigh-like term dominates: λ_sca(550) = 1/(0.0017 + 1.34c_s + 0.312c_l)
    c = 0.0075
    assert pp.get_scattering_length(550.0, m) == pytest.approx(1 / (0.0017 + 1.34 * c + 0.312 * c))


def test_frank_tamm_constant_index_matches_inverse_square():
    # With a wavelength-independent n, Frank-Tamm is ∝ 1/λ², whose inverse CDF is
    # closed form; the reference's knot construction (src/spectrum.jl:54-103)
    # must reproduce it.
    class ConstN(pp.medium.MediumProperties):
        def refractive_index(self, w):
            return np.full_like(np.asarray(w, dtype=float), 1.33)

    spec = pp.CherenkovSpectrum((300.0, 800.0), 20, ConstN())
    u = np.linspace(0, 1, 20)
    exact = 1.0 / (1 / 300.0 - u * (1 / 300.0 - 1 / 800.0))
    np.testing.assert_allclose(spec.knots, exact, atol=0.05)
    # Frank-Tamm norm: photons per cm = 2πα(1-1/n²)(1/lo - 1/hi)·1e7
    ft = pp.frank_tamm_norm((300.0, 800.0), lambda w: np.full_like(np.asarray(w, float), 1.33))
    assert ft == pytest.approx(2 * math.pi * pp.spectrum.FINE_STRUCTURE_CONSTANT * (1 - 1 / 1.33 ** 2)
                               * (1 / 300 - 1 / 800) * 1e7, rel=1e-12)


def test_cherenkov_spectrum_reference_medium_monotone():
    m = pp.make_cascadia_medium_properties()
    spec = pp.CherenkovSpectrum((300.0, 800.0), 20, m)
    assert spec.knots[0] == pytest.approx(300.0) and spec.knots[-1] == pytest.approx(800.0)
    assert np.all(np.diff(spec.knots) > 0)
    # blue-weighted: median well below the range midpoint
    assert spec.knots[10] < 450.0


def test_request_dtype_layout():
    r = pp.make_requests(2)
    assert r.dtype.itemsize == 80
    assert r["max_steps"][0] == 10


def test_julia_struct_layouts_match_header():
    """The isbits mirrors in PhotonPropHIP.jl (not executable here) lay out
    like the C structs: each field at its natural alignment, field for field
    against the ctypes mirror of include/photonprop.h (whose layout the test
    above checks in C)."""
    jl = open(os.path.join(ROOT, "photonprop.jl_amd", "julia", "PhotonPropHIP.jl")).read()
    prim = {"UInt64": (8, 8), "Int64": (8, 8), "Float64": (8, 8), "UInt32": (4, 4),
            "Int32": (4, 4), "Float32": (4, 4), "Cint": (4, 4)}
    structs = {}
    for name, body in re.findall(r"^struct (PP\w+)\n(.*?)^end", jl, re.S | re.M):
        structs[name] = [tuple(x.strip() for x in line.split("::")) for line in body.splitlines()
                         if "::" in line]

    def layout(name):
        off, align, fields = 0, 1, []
        for fname, ty in structs[name]:
            m = re.fullmatch(r"NTuple\{(\d+),(\w+)\}", ty)
            if m:
                es, ea = prim[m.group(2)]
                size, al = es * int(m.group(1)), ea
            elif ty.startswith("Ptr{"):
                size, al = 8, 8
            elif ty in prim:
                size, al = prim[ty]
            else:
                size, al, _ = layout(ty)
            off = (off + al - 1) // al * al
            fields.append((fname, off))
            off += size
            align = max(align, al)
        return (off + align - 1) // align * align, align, fields

    pairs = {"PPMediumTables": _abi.pp_medium_tables, "PPSpectrum": _abi.pp_spectrum,
             "PPEmitter": _abi.pp_emitter, "PPSphere": _abi.pp_sphere, "PPHit": None,
             "PPRequest": None, "PPServerConfig": _abi.pp_server_config,
             "PPServerStats": _abi.pp_server_stats, "PPSummary": _abi.pp_summary}
    assert set(pairs) <= set(structs), set(pairs) - set(structs)
    for jname, cty in pairs.items():
        size, _, fields = layout(jname)
        if cty is None:  # records the Python side keeps as numpy dtypes
            continue
        assert size == ctypes.sizeof(cty), (jname, size, ctypes.sizeof(cty))
        cfields = [(f[0], getattr(cty, f[0]).offset) for f in cty._fields_]
        assert [o for _, o in fields] == [o for _, o in cfields], (jname, fields, cfields)
    hit = layout("PPHit")
    assert hit[0] == _abi.HIT_DTYPE.itemsize == 64
    assert [o for _, o in hit[2]] == [_abi.HIT_DTYPE.fields[n][1] for n in _abi.HIT_DTYPE.names]
    assert layout("PPRequest")[0] == 80
