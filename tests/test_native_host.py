"""Native host code under the compilers' sanitizers (CPU only).

The serving scheduler (photonprop.jl_amd/csrc/pp_serve_core.hpp, the code
behind pp_server_*) and the load generator (pp_loadgen.cpp) run in
tests/native/serve_sim.cpp against a simulated persistent kernel on host
threads; the CPU oracle (oracle/pp_oracle.c) runs through its edge cases in
tests/native/oracle_check.c.  Each is built here with g++/gcc plain and with
AddressSanitizer + UndefinedBehaviorSanitizer, and the scheduler with
ThreadSanitizer.  The ThreadSanitizer build uses ROCm's clang++: GCC 11's
libtsan does not intercept pthread_cond_clockwait, which libstdc++ 11 uses for
condition-variable waits on the steady clock, so with it every wait_for looks
like a mutex that is never released (false races and "double lock" reports).
A deliberate fault in each sanitizer build shows that its instrumentation is
live.  Outputs go to _build/native/.

Scenarios of the scheduler (serve_sim.cpp): input checks, zero-photon requests
and FIFO queueing on slots and hit-pool room, chunked publishing of large
requests with small ones ahead, the host's idle stop and relaunch, the
kernel's own stop after waiting with nothing published, more range entries
than twice the ring (wrap), concurrent submitters with two collectors racing
for each ticket, and the open-loop load generator.
"""
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NATIVE = os.path.join(ROOT, "tests", "native")
OUT = os.path.join(ROOT, "_build", "native")
INCLUDE = os.path.join(ROOT, "include")
CORE_SRCS = [os.path.join(NATIVE, "serve_sim.cpp"),
             os.path.join(ROOT, "photonprop.jl_amd", "csrc", "pp_loadgen.cpp")]
ORACLE_SRCS = [os.path.join(NATIVE, "oracle_check.c"), os.path.join(ROOT, "oracle", "pp_oracle.c")]
ROCM_CLANG = "/opt/rocm/lib/llvm/bin/clang++"

ASAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]
ENV = dict(os.environ,
           ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
           TSAN_OPTIONS="halt_on_error=1:exitcode=66")


def _build(name, compiler, srcs, flags):
    os.makedirs(OUT, exist_ok=True)
    exe = os.path.join(OUT, name)
    cmd = [compiler, "-O1", "-g", "-pthread", "-I" + INCLUDE] + flags + srcs + ["-o", exe]
    if srcs[0].endswith(".c"):
        cmd += ["-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, "build of %s failed:\n%s" % (name, r.stderr[-4000:])
    return exe


def _run(exe, args=(), stdin=None, timeout=600):
    return subprocess.run([exe] + list(args), input=stdin, capture_output=True, text=True,
                          timeout=timeout, env=ENV)


def _kat_lines():
    with open(os.path.join(ROOT, "tests", "golden", "philox_kat.json")) as f:
        cases = json.load(f)["cases"]
    return "".join(" ".join(str(w) for w in c["counter"] + c["key"] + c["out"]) + "\n"
                   for c in cases)


def _assert_pass(r, what):
    assert r.returncode == 0 and "PASS" in r.stdout, (
        "%s: exit %d\n%s\n%s" % (what, r.returncode, r.stdout[-3000:], r.stderr[-6000:]))


def _assert_canary(c):
    # the read past the array is reported by ASan, or first by UBSan's object-size check
    assert c.returncode != 0 and ("heap-buffer-overflow" in c.stderr or
                                  "insufficient space" in c.stderr), "sanitizer canary not detected"


def test_serving_core_plain():
    exe = _build("serve_sim", "g++", CORE_SRCS, ["-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror"])
    _assert_pass(_run(exe, ["all"]), "serve_sim")


def test_serving_core_asan_ubsan():
    exe = _build("serve_sim_asan", "g++", CORE_SRCS, ["-std=c++17"] + ASAN)
    r = _run(exe, ["all"])
    _assert_pass(r, "serve_sim under ASan+UBSan")
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    _assert_canary(_run(exe, ["canary_overflow"]))


@pytest.mark.skipif(not os.path.exists(ROCM_CLANG), reason="ROCm clang++ (ThreadSanitizer runtime) absent")
def test_serving_core_tsan():
    exe = _build("serve_sim_tsan", ROCM_CLANG, CORE_SRCS, ["-std=c++17", "-fsanitize=thread"])
    r = _run(exe, ["all"])
    _assert_pass(r, "serve_sim under TSan")
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-6000:]
    c = _run(exe, ["canary_race"])
    assert c.returncode == 66 and "data race" in c.stderr, "TSan canary not detected"


def test_oracle_asan_ubsan():
    cc = "gcc"
    assert shutil.which(cc)
    exe = _build("oracle_check_asan", cc, ORACLE_SRCS,
                 ["-std=c11", "-ffp-contract=off", "-fopenmp", "-Wall", "-Wextra"] + ASAN)
    r = _run(exe, stdin=_kat_lines())
    _assert_pass(r, "oracle under ASan+UBSan")
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    _assert_canary(_run(exe, ["canary_overflow"]))
