"""Shallow-water test case 5 on the device and the stand-alone driver
(lib/swm_driver: data/namelist, test case 5 or a restart file, the time
manager's restart / diagnostics / output cadence, fort.31 budget records,
the quick-plot netCDF file).

The device run of test case 5 on the shipped g = 6 grid meets the FP64 step
bars against the oracle (1e-12 after one step, 1e-9 after several).  The
driver's outputs (budget records, restart files, the quick-plot records)
equal, bit for bit, those of the same library driven step by step from
Python.
"""
import os
import subprocess

import numpy as np
import pytest

import swm_mi355x as S
from oracle.oracle import Oracle, grid_arrays_from

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
DRIVER = os.path.join(ROOT, "shallow-water-model_amd", "lib", "swm_driver")
FIELDS = ("h", "eta", "div", "psi", "chi", "ke", "relative", "u", "tracer")


def _rel(a, b):
    d = np.abs(a - b).max()
    s = np.abs(b).max()
    return d / s if s > 0 else d


@pytest.mark.parametrize("nsteps,tol", [(1, 1e-12), (6, 1e-9)])
def test_swtc5_device_matches_oracle(nsteps, tol):
    print(f"case swtc5 g=6 centered nsteps={nsteps}", flush=True)
    ref = np.load(os.path.join(GOLD, "swtc5_g6.npz"))
    grid = S.Grid(6, p_list=ref["p_list"])
    st = S.swtc5_state(grid, 5)
    m = S.Model(grid, ntracer=5, dt=100.0, grav=S.REF_GRAV, advection="centered")
    m.init_state(st["h"], st["u"], st["tracer"], hs=st["hs"])
    orc = Oracle(grid_arrays_from(grid), 5, 100.0, S.REF_GRAV, hs=st["hs"], scheme="centered")
    orc.init_state(st["h"], st["u"], st["tracer"])
    m.step(nsteps)
    orc.step(nsteps)
    mask = grid.array("grid_center_mask") == 1.0
    for n in ("h", "eta", "div", "psi", "chi", "ke", "relative"):
        assert _rel(m.field(n)[mask], orc.field(n)[mask]) <= tol, n
    t = m.field("tracer")
    assert _rel(t.transpose(0, 1, 3, 2)[mask], orc.field("tracer").transpose(0, 1, 3, 2)[mask]) <= tol
    assert m.last_vcycles == orc.last_vcycles


def _namelist(path, hours, restart, n_restart, n_diag):
    with open(path, "w") as f:
        f.write("0000.00   ! tau_end_day\n")
        f.write("%7.2f   ! tau_end_hour\n" % hours)
        f.write("0100.00   ! dt (seconds)\n")
        f.write("%s         ! l_restart\n" % ("T" if restart else "F"))
        f.write("F         ! l_upwind\nT         ! l_centered\nF         ! l_fct\nF         ! l_remap\n")
        f.write("%d      ! n_restart\n0003      ! n_output\n%d      ! n_diagnostics\n" % (n_restart, n_diag))


def _records(path):
    raw = open(path, "rb").read()
    out, o = [], 0
    while o < len(raw):
        n = int(np.frombuffer(raw[o:o + 4], "<i4")[0])
        out.append(np.frombuffer(raw[o + 4:o + 4 + n], "<f4").copy())
        o += n + 8
    return np.array(out)


def _run_driver(d):
    r = subprocess.run([DRIVER, "-g", "4"], cwd=str(d), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def test_driver_matches_library(tmp_path):
    print("case driver g=4 swtc5 centered", flush=True)
    (tmp_path / "data").mkdir()
    (tmp_path / "restarts").mkdir()
    # tau_end = 0.12 h - 0.001: steps of 100 s until tau > 0.119 h, i.e. 5
    _namelist(tmp_path / "data" / "namelist", 0.12, False, 2, 1)
    out = _run_driver(tmp_path)
    assert out.count("GLOBAL MIN AND MAX OF SELECT FIELDS") == 5
    assert out.count(" Total Mass :") == 6
    rec = _records(tmp_path / "fort.31")
    assert rec.shape == (6, 14)
    grid = S.Grid(4)
    st = S.swtc5_state(grid, 5)
    m = S.Model(grid, ntracer=5, dt=100.0, grav=S.REF_GRAV, advection="centered")
    m.init_state(st["h"], st["u"], st["tracer"], hs=st["hs"])
    cols = [0, 1, 2, 3, 4, 7, 10, 11]
    names = {2: S.restart_name(200), 4: S.restart_name(400)}
    qp_path = tmp_path / S.qp_name(0)
    qp_want = []
    for s in range(6):
        if s:
            m.step(1)
        b = m.budgets()
        assert np.array_equal(rec[s, cols], b[:8].astype(np.float32)), s
        if s in names:
            mine = tmp_path / f"mine_{s}"
            m.write_restart(str(mine))
            assert (tmp_path / names[s]).read_bytes() == mine.read_bytes(), s
        if s in (0, 3):   # the initial state and n_output = 3
            qp_want.append((s * 100, _qp_fields(grid, m, st["hs"])))
    _check_qp(qp_path, qp_want, 0)
    # restart from the file of step 4 (l_restart): 100 s steps from 400 s
    # until tau > 0.169 h, i.e. 3 steps
    (tmp_path / "restarts" / "swm_restart").write_bytes((tmp_path / names[4]).read_bytes())
    _namelist(tmp_path / "data" / "namelist", 0.17, True, 9999, 3)
    out = _run_driver(tmp_path)
    assert "finished reading initial conditions" in out
    rec = _records(tmp_path / "fort.31")
    assert rec.shape == (2, 14)
    m2 = S.Model(grid, ntracer=5, dt=100.0, grav=S.REF_GRAV, advection="centered")
    m2.init_state(st["h"], st["u"], st["tracer"], hs=st["hs"])
    m2.read_restart(str(tmp_path / names[4]))
    assert np.array_equal(rec[0, cols], m2.budgets()[:8].astype(np.float32))
    qp_want = [(400, _qp_fields(grid, m2, st["hs"]))]
    m2.step(3)
    assert m2.time_seconds == 700
    assert np.array_equal(rec[1, cols], m2.budgets()[:8].astype(np.float32))
    qp_want.append((700, _qp_fields(grid, m2, st["hs"])))
    # the file is named for model time 0 and its first period starts there
    # (initialize.F:153-159 runs before the restart read)
    _check_qp(qp_path, qp_want, 0)


def _qp_fields(grid, m, hs):
    mask = (grid.array("grid_center_mask") > 0.5).ravel(order="F")
    pk = lambda x: np.asarray(x).ravel(order="F")[mask].astype(np.float32)
    h = m.field("h")
    t = m.field("tracer")
    out = {"thickness": pk(h), "height": pk(h + hs), "relative": pk(m.field("relative")),
           "divergence": pk(m.field("div")), "absolute": pk(m.field("eta")), "kinetic_energy": pk(m.field("ke"))}
    for k in range(5):
        out[f"tracer_{k + 1}"] = pk(t[:, :, k, :])
    return out


def _check_qp(path, want, t0):
    from scipy.io import netcdf_file
    nc = netcdf_file(str(path), "r", mmap=False)
    v = nc.variables
    assert v["time"].shape[0] == len(want)
    start = t0
    for r, (t, f) in enumerate(want):
        assert v["start_period"][r] == start / 86400.0 and v["end_period"][r] == t / 86400.0, r
        for name, x in f.items():
            assert np.array_equal(v[name][r], x), (r, name)
        start = t
    nc.close()
