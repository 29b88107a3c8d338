"""Device time step against the oracle (the CPU FP64 restatement of the
reference step, itself pinned bit for bit to the compiled reference).

The device path computes the same operations in the same order except for
the global sums of the elliptic solve (mean removal, residual norm), which
are reduced in a different order; the FP64 bar is therefore a relative
difference of 1e-12 after one step and 1e-9 after a run of steps, on every
prognostic field at real cells and corners.
"""
import numpy as np
import pytest

import swm_mi355x as S
from swm_mi355x import ics
from oracle.oracle import Oracle, grid_arrays_from

pytestmark = pytest.mark.gpu

TOL_1 = 1e-12
TOL_N = 1e-9


def _rel(a, b, m):
    d = np.abs(a - b)[m].max()
    s = np.abs(b)[m].max()
    return d / s if s > 0 else d


def _compare(model, orc, grid, tol):
    mask = grid.array("grid_center_mask") == 1.0
    cm = grid.array("grid_corner_mask") == 1.0
    out = {}
    for name in ("h", "eta", "div", "psi", "chi", "ke", "relative"):
        out[name] = _rel(model.field(name), orc.field(name), mask)
    u = model.field("u")
    out["u"] = _rel(u, orc.field("u"), np.broadcast_to(cm[None], u.shape))
    t = model.field("tracer")
    out["tracer"] = _rel(t, orc.field("tracer"), np.broadcast_to(mask[:, :, None, :], t.shape))
    bad = {k: v for k, v in out.items() if not v <= tol}
    assert not bad, f"relative differences above {tol}: {bad} (all: {out})"
    return out


@pytest.mark.parametrize("g,nsteps", [(3, 1), (4, 1), (5, 1), (4, 12), (5, 6)])
def test_step_matches_oracle(g, nsteps):
    print(f"case g={g} nsteps={nsteps}", flush=True)
    grid = S.Grid(g)
    dt = 450.0 * 2.0 ** (5 - g)
    st = ics.synthetic_state(grid, ntracer=2, lmax=10)
    model = S.Model(grid, ntracer=2, dt=dt, grav=S.REF_GRAV)
    model.init_state(st["h"], st["u"], st["tracer"], hs=st["hs"])
    orc = Oracle(grid_arrays_from(grid), 2, dt, S.REF_GRAV)
    orc.init_state(st["h"], st["u"], st["tracer"])
    _compare(model, orc, grid, 1e-14)
    for _ in range(nsteps):
        model.step(1)
        orc.step(1)
    _compare(model, orc, grid, TOL_1 if nsteps == 1 else TOL_N)
    assert model.last_vcycles == orc.last_vcycles


def test_mass_conserved():
    grid = S.Grid(4)
    st = ics.synthetic_state(grid, ntracer=1, lmax=8)
    model = S.Model(grid, ntracer=1, dt=900.0, grav=S.REF_GRAV)
    model.init_state(st["h"], st["u"], st["tracer"], hs=st["hs"])
    m0 = model.total_mass()
    model.step(20)
    m1 = model.total_mass()
    assert abs(m1 - m0) / m0 < 1e-13


@pytest.mark.parametrize("scheme,g,nsteps", [("centered", 4, 1), ("centered", 4, 12), ("upwind", 4, 1),
                                             ("upwind", 5, 8), ("centered", 5, 8)])
def test_advection_schemes_match_oracle(scheme, g, nsteps):
    # step_dt with l_centered / l_upwind (step.F:221-327, :495-602); the oracle
    # is pinned bit for bit to the compiled reference for both (test_golden.py)
    print(f"case {scheme} g={g} nsteps={nsteps}", flush=True)
    grid = S.Grid(g)
    dt = 200.0 * 2.0 ** (5 - g)
    st = ics.synthetic_state(grid, ntracer=2, lmax=10)
    model = S.Model(grid, ntracer=2, dt=dt, grav=S.REF_GRAV, advection=scheme)
    model.init_state(st["h"], st["u"], st["tracer"], hs=st["hs"])
    orc = Oracle(grid_arrays_from(grid), 2, dt, S.REF_GRAV, scheme=scheme)
    orc.init_state(st["h"], st["u"], st["tracer"])
    model.step(nsteps)
    orc.step(nsteps)
    _compare(model, orc, grid, TOL_1 if nsteps == 1 else TOL_N)
    assert model.last_vcycles == orc.last_vcycles


@pytest.mark.parametrize("scheme", ["centered", "upwind"])
def test_advection_schemes_mixed_close_to_fp64(scheme):
    grid = S.Grid(5)
    st = ics.synthetic_state(grid, ntracer=1, lmax=10)
    out = {}
    for prec in (S.SWM_FP64, S.SWM_MIXED):
        m = S.Model(grid, ntracer=1, dt=200.0, grav=S.REF_GRAV, precision=prec, advection=scheme)
        m.init_state(st["h"], st["u"], st["tracer"], hs=st["hs"])
        m.step(20)
        out[prec] = {n: m.field(n) for n in ("h", "eta", "tracer")}
    mask = grid.array("grid_center_mask") == 1.0
    for n in ("h", "eta"):
        a, b = out[S.SWM_FP64][n], out[S.SWM_MIXED][n]
        assert _rel(b, a, mask) <= 1e-6, n
    a, b = out[S.SWM_FP64]["tracer"], out[S.SWM_MIXED]["tracer"]
    assert _rel(b, a, np.broadcast_to(mask[:, :, None, :], a.shape)) <= 1e-5
