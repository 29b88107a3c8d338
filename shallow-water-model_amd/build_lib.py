"""Builds the gfx950 shared library of the dynamical core in-tree.

    python shallow-water-model_amd/build_lib.py

Host set-up code is compiled with g++ and the device code with hipcc for
gfx950; both without floating-point contraction, so that the FP64 path keeps
the reference's operation order.  The Fortran modules (fortran/swm_mi355x.F90,
ISO_C_BINDING over the C ABI) and the Fortran driver check program are
compiled with ROCm's flang.  Output under shallow-water-model_amd/lib/:
libswm_mi355x.so, swm_fortran_check, the stand-alone driver swm_driver
(driver/swm_driver.cpp) and the Fortran module files in obj/.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(OUT_DIR, "libswm_mi355x.so")
FORTRAN_CHECK = os.path.join(OUT_DIR, "swm_fortran_check")
DRIVER = os.path.join(OUT_DIR, "swm_driver")

HOST_SOURCES = ["swm_setup.cpp", "swm_metrics.cpp", "swm_remap_setup.cpp", "swm_diss_setup.cpp", "swm_capi_grid.cpp",
                "swm_restart.cpp", "swm_qp.cpp", "swm_ics.cpp", "swm_namelist.cpp"]
DEVICE_SOURCES = ["swm_device.hip"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the library needs ROCm's hipcc")


def _flang():
    for cand in (os.environ.get("FLANG"), "/opt/rocm/lib/llvm/bin/flang", shutil.which("flang")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("flang not found: the Fortran modules need ROCm's flang")


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
        # code-object v5: the kernels read their workgroup size from the
        # implicit kernel arguments at the offset of that layout
        _run([hipcc, "-std=c++17", "-O3", "-fPIC", "--offload-arch=gfx950", "-mcode-object-version=5",
              "-ffp-contract=off",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-Wno-unused-result", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # Fortran modules and the driver check program (runs against the .so
    # beside it)
    flang = _flang()
    fdir = os.path.join(HERE, "fortran")
    fobj = os.path.join(obj_dir, "swm_mi355x.F90.o")
    _run([flang, "-O2", "-module-dir", obj_dir, "-c", os.path.join(fdir, "swm_mi355x.F90"), "-o", fobj])
    tmp = FORTRAN_CHECK + ".tmp"
    _run([flang, "-O2", "-module-dir", obj_dir, os.path.join(fdir, "swm_fortran_check.F90"), fobj,
          "-L" + OUT_DIR, "-lswm_mi355x", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, FORTRAN_CHECK)
    # the stand-alone driver program
    tmp = DRIVER + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-Wall", os.path.join(HERE, "driver", "swm_driver.cpp"),
          "-L" + OUT_DIR, "-lswm_mi355x", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, DRIVER)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
