"""Build libphotonprop.so for gfx950 in-tree (photonprop.jl_amd/lib/), and the
host-only serving load generator libpploadgen.so beside it.

    python photonprop.jl_amd/build_lib.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = [os.path.join(HERE, "csrc", "pp_kernels.hip"), os.path.join(HERE, "csrc", "pp_mlp.hip")]
OUT = os.path.join(HERE, "lib", "libphotonprop.so")
LOADGEN_SRC = os.path.join(HERE, "csrc", "pp_loadgen.cpp")
LOADGEN_OUT = os.path.join(HERE, "lib", "libpploadgen.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # fused multiply-adds only where the source writes them
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    # HSA runtime: result copies on an SDMA engine (serving)
    "-lhsa-runtime64",
]


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = ["hipcc"] + HIPCC_FLAGS + ["-o", tmp] + SRC
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed for libphotonprop.so")
    if verbose:
        sys.stderr.write(res.stderr)
    os.replace(tmp, OUT)
    build_loadgen()
    return OUT


def build_loadgen():
    """The open-loop load generator: plain C++ over the public header, no HIP."""
    tmp = LOADGEN_OUT + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", INCLUDE, "-o", tmp,
           LOADGEN_SRC, "-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed for libpploadgen.so")
    os.replace(tmp, LOADGEN_OUT)
    return LOADGEN_OUT


if __name__ == "__main__":
    print(build(verbose=True))
