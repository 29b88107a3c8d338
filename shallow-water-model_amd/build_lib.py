"""Builds the gfx950 shared library of the dynamical core in-tree.

    python shallow-water-model_amd/build_lib.py

Host set-up code is compiled with g++ and the device code with hipcc for
gfx950; both without floating-point contraction, so that the FP64 path keeps
the reference's operation order.  Output: shallow-water-model_amd/lib/
libswm_mi355x.so.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(OUT_DIR, "libswm_mi355x.so")

HOST_SOURCES = ["swm_setup.cpp", "swm_metrics.cpp", "swm_capi_grid.cpp"]
DEVICE_SOURCES = ["swm_device.hip"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the library needs ROCm's hipcc")


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(OUT_DIR, exist_ok=True)
    obj_dir = os.path.join(OUT_DIR, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(obj_dir, src + ".o")
        _run(["g++", "-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
              "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    hipcc = _hipcc()
    for src in DEVICE_SOURCES:
        obj = os.path.join(obj_dir, src + ".o")
        _run([hipcc, "-std=c++17", "-O3", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-Wno-unused-result", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
