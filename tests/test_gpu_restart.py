"""Restart of the device model (swm_write_restart / swm_read_restart, after
swm/source/swm_restart.F) against the compiled reference's runs in
restart_g3.npz (tests/golden/make_golden.py) and against the oracle.

The device continues from a restart file as the reference did: the state the
reference derived after read_swm_restart within 1e-14 (relative), then its
steps within the FP64 step bars (1e-12 after one step, 1e-9 after several).
The files the device writes hold the state of the real cells exactly.  All
runs are over a mountain (non-zero surface height).
"""
import os

import numpy as np
import pytest

import swm_mi355x as S
from swm_mi355x import ics
from oracle.oracle import Oracle, grid_arrays_from

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NT = 5
TOL_0, TOL_1, TOL_N = 1e-14, 1e-12, 1e-9
FIELDS = ("h", "eta", "div", "psi", "chi", "ke", "relative", "u", "tracer")


@pytest.fixture(scope="module")
def ref():
    return np.load(os.path.join(GOLD, "restart_g3.npz"))


@pytest.fixture(scope="module")
def grid():
    return S.Grid(3)


def _rel(a, b):
    d = np.abs(a - b).max()
    s = np.abs(b).max()
    return d / s if s > 0 else d


def _compare(model, ref, prefix, tol):
    out = {n: _rel(model.field(n).reshape(-1, order="F"), ref[f"{prefix}_{n}"]) for n in FIELDS}
    bad = {k: v for k, v in out.items() if not v <= tol}
    assert not bad, f"{prefix}: relative differences above {tol}: {bad}"


def _model(grid, ref, precision=S.SWM_FP64):
    m = S.Model(grid, ntracer=NT, dt=float(ref["dt"]), grav=S.REF_GRAV, precision=precision)
    # the surface height of the run; the restart replaces h, u and tracers
    m.init_state(ref["ic_h"], ref["ic_u"], ref["ic_tracer"], hs=ref["ic_hs"])
    return m


def test_device_continues_from_reference_file(ref, grid, tmp_path):
    print("case restart: reference file, 4 steps", flush=True)
    p = tmp_path / "swm_restart"
    p.write_bytes(bytes(ref["a_file"]))
    m = _model(grid, ref)
    m.step(2)   # state, clock and step counter are all replaced
    m.read_restart(str(p))
    assert m.time_seconds == 2400 and m.step_count == 0
    _compare(m, ref, "b_s0", TOL_0)
    m.step(1)
    _compare(m, ref, "b_s1", TOL_1)
    m.step(3)
    _compare(m, ref, "b_s4", TOL_N)
    assert m.time_seconds == 2400 + 4 * 800


def test_device_continues_from_library_file_and_writes_reference_name(ref, grid, tmp_path, monkeypatch):
    print("case restart: library file, 2 steps, default names", flush=True)
    monkeypatch.chdir(tmp_path)
    os.mkdir("restarts")
    S.write_restart_file(grid, "restarts/swm_restart", ref["c_in_h"], ref["c_in_u"], ref["c_in_tracer"],
                         int(ref["c_ts"]))
    m = _model(grid, ref)
    m.read_restart()   # restarts/swm_restart
    _compare(m, ref, "c_s0", TOL_0)
    m.step(2)
    _compare(m, ref, "c_s2", TOL_N)
    m.write_restart()   # restarts/swm_restart.dDDDDDD.hh:mm:ss of the model time
    assert os.path.exists(str(ref["c_name"]))
    ts, h, u, tr = S.read_restart_file(grid, str(ref["c_name"]), NT)
    raw = tmp_path / "theirs"
    raw.write_bytes(bytes(ref["c_file"]))
    ts_ref, h_ref, u_ref, tr_ref = S.read_restart_file(grid, str(raw), NT)
    assert ts == ts_ref == int(ref["c_ts"]) + 1600
    for a, b in ((h, h_ref), (u, u_ref), (tr, tr_ref)):
        assert _rel(a, b) <= TOL_N


@pytest.mark.parametrize("precision", [S.SWM_FP64, S.SWM_MIXED])
def test_write_read_round_trip_then_oracle(precision, tmp_path):
    print(f"case restart round trip precision={precision}", flush=True)
    g, nt, dt = 4, 2, 400.0
    grid = S.Grid(g)
    st = ics.synthetic_state(grid, ntracer=nt, lmax=10, seed_index=7)
    m = S.Model(grid, ntracer=nt, dt=dt, grav=S.REF_GRAV, precision=precision)
    m.init_state(st["h"], st["u"], st["tracer"], hs=st["hs"])
    m.step(5)
    p = str(tmp_path / "r")
    m.write_restart(p)
    ts, h, u, tr = S.read_restart_file(grid, p, nt)
    assert ts == m.time_seconds == 5 * 400
    mask = grid.array("grid_center_mask") == 1.0
    # the file holds exactly what swm_get_field returns, at the real cells
    assert np.array_equal(h[mask], m.field("h")[mask])
    assert np.array_equal(u[:, :, mask], m.field("u")[:, :, mask])
    assert np.array_equal(tr.transpose(0, 1, 3, 2)[mask], m.field("tracer").transpose(0, 1, 3, 2)[mask])
    # a second context reads it and continues; the oracle from the same arrays
    m2 = S.Model(grid, ntracer=nt, dt=dt, grav=S.REF_GRAV, precision=precision)
    m2.init_state(st["h"], st["u"], st["tracer"], hs=st["hs"])
    m2.read_restart(p)
    assert m2.time_seconds == 2000
    orc = Oracle(grid_arrays_from(grid), nt, dt, S.REF_GRAV, hs=st["hs"])
    orc.init_state(h, u, tr)
    m2.step(4)
    orc.step(4)
    # the mixed path: its FP32 storage and its own elliptic schedule (the
    # solves stop at other residuals within the criterion), which the small
    # divergence field shows most: h 1e-6, eta and u 1e-5, div 1e-4
    mixed_tol = {"h": 1e-6, "eta": 1e-5, "div": 1e-4, "u": 1e-5}
    for n in ("h", "eta", "div", "u"):
        tol = TOL_N if precision == S.SWM_FP64 else mixed_tol[n]
        a, b = m2.field(n), orc.field(n)
        if n == "u":
            a, b = a[:, :, mask], b[:, :, mask]
        else:
            a, b = a[mask], b[mask]
        assert _rel(a, b) <= tol, n


def test_clock_truncates_like_the_reference():
    # time_manager_module.F:96: nsecond_dyn = nsecond_dyn + dt, an integer
    # (truncated) after each step; 12.5 s steps advance it by 12 s
    grid = S.Grid(3)
    st = ics.synthetic_state(grid, ntracer=1, lmax=4)
    m = S.Model(grid, ntracer=1, dt=12.5, grav=S.REF_GRAV)
    m.init_state(st["h"], st["u"], st["tracer"], hs=st["hs"])
    assert m.time_seconds == 0
    m.step(4)
    assert m.time_seconds == 48


def test_restart_errors(ref, grid, tmp_path):
    m = _model(grid, ref)
    with pytest.raises(S.SwmError):
        m.read_restart(str(tmp_path / "missing"))
    other = S.Model(grid, ntracer=NT - 1, dt=800.0, grav=S.REF_GRAV)
    p = tmp_path / "f"
    p.write_bytes(bytes(ref["a_file"]))
    with pytest.raises(S.SwmError) as e:
        other.read_restart(str(p))
    assert e.value.status == S.SWM_ERR_DIMS
    fresh = S.Model(grid, ntracer=NT, dt=800.0, grav=S.REF_GRAV)
    with pytest.raises(S.SwmError) as e:
        fresh.write_restart(str(tmp_path / "w"))
    assert e.value.status == S.SWM_ERR_STATE
