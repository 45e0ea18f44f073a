"""Build libparacfd.so in-tree: host preprocessing (C++20) + HIP kernels for gfx950."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libparacfd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PCFD_OFFLOAD_ARCH", "gfx950")

# packed FP32 instructions are disabled for all device code: some of their forms have given
# wrong results on this part now and then, with no error
HIP_FLAGS = ["-std=c++20", "-O3", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-munsafe-fp-atomics", "-Wno-unused-result"]
# no fused multiply-add in the host preprocessing or the device build: the two evaluate the
# shared per-cell code (solid.hpp, build_graph.hpp) with the same IEEE operations and agree bit
# for bit
HOST_FLAGS = ["-std=c++20", "-O2", "-fPIC", "-ffp-contract=off"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in ("topology.cpp", "nonorth.cpp", "capi_topology.cpp"):
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        _run(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    obj = os.path.join(BUILD, "solver.o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "solver.hip"), "-o", obj])
    objs.append(obj)
    for src in ("build_device.hip", "build_multigrid.hip"):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-ffp-contract=off", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = OUT + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
