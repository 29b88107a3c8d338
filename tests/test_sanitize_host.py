"""The host half of the library (grid set-up, halo tables, metrics,
multigrid weights, restart codec, test case 5, namelist reader) and the
oracle, built with AddressSanitizer and UndefinedBehaviorSanitizer (g++,
host code only) and run through tests/sanitize/host_check.cpp: no memory
error and no undefined behaviour at grid levels 1-5 and in three oracle
runs at level 3 (one per advection scheme)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "shallow-water-model_amd", "csrc")
HOST = ["swm_setup.cpp", "swm_metrics.cpp", "swm_remap_setup.cpp", "swm_capi_grid.cpp", "swm_restart.cpp",
        "swm_ics.cpp", "swm_namelist.cpp"]


@pytest.mark.timeout(600)
def test_host_code_under_sanitizers(tmp_path):
    if shutil.which("g++") is None or shutil.which("gcc") is None:
        pytest.skip("no host compiler")
    flags = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
             "-ffp-contract=off"]
    objs = []
    for src in HOST:
        obj = str(tmp_path / (src + ".o"))
        subprocess.run(["g++", "-std=c++17", *flags, "-c", os.path.join(CSRC, src), "-o", obj], check=True)
        objs.append(obj)
    orc = str(tmp_path / "oracle.o")
    subprocess.run(["gcc", "-std=c11", *flags, "-c", os.path.join(ROOT, "oracle", "swm_oracle.c"), "-o", orc],
                   check=True)
    drv = str(tmp_path / "host_check.o")
    subprocess.run(["g++", "-std=c++17", *flags, "-c", os.path.join(ROOT, "tests", "sanitize", "host_check.cpp"),
                    "-o", drv], check=True)
    exe = str(tmp_path / "host_check")
    subprocess.run(["g++", *flags, drv, orc, *objs, "-o", exe, "-lm"], check=True)
    # the environment's preloaded libraries stay as they are; ASan is told
    # not to insist on coming first
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=500, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "host_check ok" in r.stdout
