"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer: the topology builders
(AMR, cut cells, compact edges, wall moments, multigrid, dense coarse inverse, batches, the
non-orthogonal records), the topology C ABI and the FP64 oracle, compiled with g++ and driven by
tests/native/sanitize_main.cpp.  Any sanitizer report fails the run (-fno-sanitize-recover)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "paracfd_amd", "csrc")
SOURCES = [os.path.join(CSRC, f) for f in ("topology.cpp", "nonorth.cpp", "capi_topology.cpp")]
DRIVER = os.path.join(ROOT, "tests", "native", "sanitize_main.cpp")
BINARY = os.path.join(ROOT, "paracfd_amd", "_build", "sanitize_main")


def _stale():
    if not os.path.exists(BINARY):
        return True
    deps = SOURCES + [DRIVER, os.path.join(ROOT, "oracle", "oracle.cpp"), os.path.join(CSRC, "topology.hpp"),
                      os.path.join(CSRC, "lsq3.hpp"), os.path.join(ROOT, "include", "paracfd.h")]
    return max(os.path.getmtime(p) for p in deps) > os.path.getmtime(BINARY)


def test_host_code_under_asan_and_ubsan():
    os.makedirs(os.path.dirname(BINARY), exist_ok=True)
    if _stale():
        cmd = ["g++", "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
               "-fno-sanitize-recover=undefined", *SOURCES, DRIVER, "-o", BINARY, "-lpthread"]
        subprocess.run(cmd, check=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([BINARY], capture_output=True, text=True, env=env, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "sanitize ok" in r.stdout
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
