"""Resource limits of the built gfx950 kernels, read from the code object's
metadata (CPU only: the library built in-tree, unbundled with the ROCm LLVM
tools, no GPU needed).

propagate_kernel is bound by VALU issue at 8 waves per SIMD: that needs at most
64 VGPRs, and a spilled VGPR costs scratch traffic and instructions.  The serve
kernel shares its step loop and must keep the same occupancy.
"""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(os.path.dirname(HERE), "photonprop.jl_amd", "lib", "libphotonprop.so")
LLVM = "/opt/rocm/llvm/bin"
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"


def _tool(name):
    p = os.path.join(LLVM, name)
    return p if os.path.exists(p) else shutil.which(name)


@pytest.fixture(scope="module")
def kernels():
    tools = [_tool(n) for n in ("llvm-objcopy", "clang-offload-bundler", "llvm-readelf")]
    if not os.path.exists(LIB):
        pytest.skip("libphotonprop.so not built")
    if not all(tools):
        pytest.skip("ROCm LLVM tools not found")
    objcopy, bundler, readelf = tools
    notes = ""
    with tempfile.TemporaryDirectory() as d:
        fb = os.path.join(d, "fatbin")
        subprocess.run([objcopy, "--dump-section", ".hip_fatbin=" + fb, LIB], check=True,
                       capture_output=True)
        # one offload bundle per translation unit, concatenated in the section
        data = open(fb, "rb").read()
        magic = b"__CLANG_OFFLOAD_BUNDLE__"
        starts = [m.start() for m in re.finditer(re.escape(magic), data)]
        assert starts, "no offload bundle in " + LIB
        for i, a in enumerate(starts):
            part, co = os.path.join(d, "b%d" % i), os.path.join(d, "b%d.elf" % i)
            with open(part, "wb") as f:
                f.write(data[a:starts[i + 1] if i + 1 < len(starts) else len(data)])
            subprocess.run([bundler, "--unbundle", "--type=o", "--input=" + part,
                            "--targets=" + TARGET, "--output=" + co], check=True,
                           capture_output=True)
            notes += subprocess.run([readelf, "--notes", co], check=True, capture_output=True,
                                    text=True).stdout
    # kernels are the items of the top-level list (indent 2, "- ."); the
    # deeper items are their arguments
    entries, cur = [], None
    for line in notes.splitlines():
        m = re.match(r"( *)(- +)?\.(\w+):\s+(\S+)", line)
        if not m:
            continue
        indent, dash, key, val = m.groups()
        if dash and len(indent) == 2:
            cur = {}
            entries.append(cur)
        if cur is None or len(indent) + len(dash or "") != 4:
            continue
        cur[key] = val if key == "name" else (int(val) if val.isdigit() else val)
    return {e["name"]: e for e in entries if str(e.get("name", "")).startswith("_Z")}


def _find(kernels, fragment):
    hits = {k: v for k, v in kernels.items() if fragment in k}
    assert hits, "kernel %s not in the code object" % fragment
    return hits


def test_propagate_kernel_fits_eight_waves_without_spills(kernels):
    for name, md in _find(kernels, "propagate_kernel").items():
        assert md["vgpr_count"] <= 64, (name, md)
        assert md["vgpr_spill_count"] == 0, (name, md)
        assert md["private_segment_fixed_size"] == 0, (name, md)


def test_serve_kernel_keeps_the_batch_kernels_occupancy(kernels):
    for name, md in _find(kernels, "serve_kernel").items():
        assert md["vgpr_count"] <= 64, (name, md)
        assert md["vgpr_spill_count"] == 0, (name, md)
        assert md["private_segment_fixed_size"] == 0, (name, md)


def test_fused_mlp_forward_without_spills(kernels):
    for name, md in _find(kernels, "mlp_forward_fused_kernel").items():
        assert md["vgpr_spill_count"] == 0, (name, md)
        assert md["private_segment_fixed_size"] == 0, (name, md)
