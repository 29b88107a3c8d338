"""The C ABI library loads and exports every function include/swm_mi355x.h
declares; host-only entry points work without a device and device entry
points fail loudly (no CPU fallback)."""
import ctypes
import os
import re

import numpy as np
import pytest

import swm_mi355x as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "swm_mi355x.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(swm_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) > 20
    lib = ctypes.CDLL(S.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_grid_errors_are_reported():
    with pytest.raises(S.SwmError) as e:
        S.Grid(0)
    assert e.value.status == 1
    with pytest.raises(KeyError):
        S.Grid(2).array("no_such_array")


def test_no_device_is_an_error_not_a_fallback():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(S.SwmError) as e:
        S.Model(S.Grid(2))
    assert e.value.status == 6


def test_grid_points_are_unit_vectors_in_reference_order():
    p = S.grid_points(4)
    assert p.shape == (3, 2562)
    assert np.allclose(np.linalg.norm(p, axis=0), 1.0, atol=1e-15)
    assert np.array_equal(p[:, 0], [0.0, 0.0, 1.0]) and np.array_equal(p[:, 1], [0.0, 0.0, -1.0])


def test_fortran_check_program_reports_missing_device(tmp_path):
    # without a GPU the Fortran driver stops at swm_create with the library's
    # message, as the reference stops through abort_code
    import subprocess
    exe = os.path.join(ROOT, "shallow-water-model_amd", "lib", "swm_fortran_check")
    assert os.access(exe, os.X_OK)
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    import numpy as np
    g, nt = 3, 1
    n = 2 + 2 ** g
    fin = tmp_path / "in.bin"
    with open(fin, "wb") as f:
        f.write(np.array([g, nt, 1], dtype="<i4").tobytes())
        f.write(np.array([900.0, 9.8, 6.37e6, 7.292e-5], dtype="<f8").tobytes())
        f.write(np.zeros(n * n * 10 * (2 + 4 + nt)).tobytes())
    r = subprocess.run([exe, str(fin), str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "swm_create failed, status 6" in r.stdout
