"""Incremental remapping on the device (SWM_ADV_REMAP: get_remap_flux,
swm/source/step.F:687-771, over utilities/layer_2/remap.F).

  - the compiled reference's runs (tests/golden/run_g<g>_remap.npz, g = 3
    and 4, 12 steps) with the tracer limiter switches each reference build
    used (it leaves them without a value): after the first step the
    remapped thickness and tracers and the divergence and relative
    vorticity bit for bit; the fields of the elliptic solve (psi, chi,
    absolute vorticity, velocity, kinetic energy) to 1e-12 of their
    largest value, as the FP64 step (the device's global sums of the solve
    add in another order than the reference's sequential loop); after 2, 3
    and 12 steps every field to 1e-10, 1e-10 and 1e-9;
  - the oracle over 24 steps of 100 s at level 6 with the default two-sided
    limiter
    to the same bars, and remap_mass_and_tracers alone
    (swm_op_remap_mass_and_tracers) bit for bit at level 5 with each limiter
    setting;
  - at level 8 (655362 cells, dt = 25 s, 48 steps): the remapped thickness
    and mass-weighted tracer are conserved in the X3 cell areas the
    remapping's flux divergence uses, to 1e-12 relative (the ABM3 sums of
    conservative increments);
  - a thickness below -1e-10 entering the remapping stops the step with
    SWM_ERR_STATE (the reference stops in remap_prep), and the contexts the
    remapping does not run in are refused.
"""
import os

import numpy as np
import pytest

import swm_mi355x as S
from swm_mi355x import ics
from oracle.oracle import Oracle, grid_arrays_from, remap_arrays_from

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("h", "eta", "div", "psi", "chi", "ke", "relative", "u", "tracer")
# switches of the reference builds (see tests/test_golden.py)
LIMITER = {3: lambda step: (1, 1) if step == 1 else (0, 0), 4: lambda step: (1, 1)}
EXACT_STEP1 = ("h", "tracer", "div", "relative")
BAR = {1: 1e-12, 2: 1e-10, 3: 1e-10, 12: 1e-9, 24: 1e-9}


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _grid(g):
    if g == 4:
        return S.Grid(4, p_list=np.load(os.path.join(GOLD, "setup_g4.npz"))["p_list"])
    return S.Grid(g)


@pytest.mark.parametrize("g", [3, 4])
def test_remap_device_matches_reference(g):
    print(f"case remap device vs reference g={g}", flush=True)
    ref = np.load(os.path.join(GOLD, f"run_g{g}_remap.npz"))
    grid = _grid(g)
    m = S.Model(grid, ntracer=5, dt=float(ref["dt"]), grav=S.REF_GRAV, advection="remap")
    m.init_state(ref["ic_h"], ref["ic_u"], ref["ic_tracer"], hs=ref["ic_hs"])
    done = 0
    for s in (1, 2, 3, int(ref["nsteps"])):
        while done < s:
            done += 1
            m.set_remap_limiter(*LIMITER[g](done))
            m.step(1)
        for name in FIELDS:
            mine = m.field(name).reshape(-1, order="F")
            want = ref[f"s{s}_{name}"]
            if s == 1 and name in EXACT_STEP1:
                assert np.array_equal(mine, want), (g, s, name)
            else:
                assert _rel(mine, want) <= BAR[s], (g, s, name, _rel(mine, want))
    m.close()


def test_remap_device_matches_oracle_level6():
    print("case remap device vs oracle g=6, 24 steps", flush=True)
    grid = S.Grid(6)
    st = ics.synthetic_state(grid, ntracer=3, lmax=20, seed_index=6)
    dt = 100.0
    m = S.Model(grid, ntracer=3, dt=dt, grav=S.REF_GRAV, advection="remap")
    m.init_state(st["h"], st["u"], st["tracer"], hs=st["hs"])
    orc = Oracle(grid_arrays_from(grid), 3, dt, S.REF_GRAV, hs=st["hs"], scheme="remap",
                 remap=remap_arrays_from(grid))
    orc.init_state(st["h"], st["u"], st["tracer"])
    for n in (1, 1, 1, 21):
        m.step(n)
        orc.step(n)
        for name in FIELDS:
            a, b = m.field(name), orc.field(name)
            if m.step_count == 1 and name in EXACT_STEP1:
                assert np.array_equal(a, b), (m.step_count, name)
            else:
                assert _rel(a, b) <= BAR[m.step_count], (m.step_count, name, _rel(a, b))
    m.close()


@pytest.mark.parametrize("limiter", [(1, 1), (0, 0), (1, 0)])
def test_remap_operator_matches_oracle(limiter):
    """remap_mass_and_tracers alone (swm_op_remap_mass_and_tracers) against
    the oracle bit for bit at level 5: masses below eps (set to zero with
    their tracers), corners whose first velocity component is zero (the
    reference's speed test then keeps the corner fixed), each limiter
    setting; a mass below -eps is refused."""
    print(f"case remap operator g=5 limiter={limiter}", flush=True)
    grid = S.Grid(5)
    st = ics.synthetic_state(grid, ntracer=3, lmax=20, seed_index=5)
    dt = 600.0
    mass = np.array(st["h"], order="F")
    mass[3, 3, 2] = 5e-11
    mass[4, 7, 5] = 0.0
    trace = np.array(st["tracer"], order="F")
    uc = np.array(st["u"], order="F")
    uc[0, :, 5, 5, 1] = 0.0
    uc[0, 1, 9, 4, 7] = 0.0
    m = S.Model(grid, ntracer=3, dt=dt, grav=S.REF_GRAV, advection="remap")
    m.set_remap_limiter(*limiter)
    a = m.remap_mass_and_tracers(mass, trace, dt, uc)
    orc = Oracle(grid_arrays_from(grid), 3, dt, S.REF_GRAV, scheme="remap", remap=remap_arrays_from(grid),
                 limiter=limiter)
    b = orc.remap_mass_and_tracers(mass, trace, uc)
    for name, x, y in zip(("mass", "trace", "edgeflux"), a, b):
        assert np.array_equal(x, y), name
    # the remapping moves mass without creating any: sums over the real cells
    # in the X3 areas of its flux divergence, the input after remap_prep
    real = grid.array("grid_center_mask") == 1.0
    ax3 = grid.array("remap_area_x3")
    before = np.sum((ax3 * np.where(mass < 1e-10, 0.0, mass))[real])
    assert abs(np.sum((ax3 * a[0])[real]) - before) <= 1e-13 * before
    bad = mass.copy()
    bad[6, 6, 3] = -1.0
    with pytest.raises(S.SwmError, match="mass below"):
        m.remap_mass_and_tracers(bad, trace, dt, uc)
    m.close()
    orc.close()


def test_remap_conserves_mass_level8():
    print("case remap g=8 conservation, 48 steps", flush=True)
    grid = S.Grid(8)
    st = ics.synthetic_state(grid, ntracer=2, lmax=40, seed_index=8)
    m = S.Model(grid, ntracer=2, dt=25.0, grav=S.REF_GRAV, advection="remap")
    m.init_state(st["h"], st["u"], st["tracer"], hs=st["hs"])
    mask = grid.array("grid_center_mask") == 1.0
    ax3 = grid.array("remap_area_x3")

    def sums():
        h = m.field("h")
        t = m.field("tracer")
        return np.sum((ax3 * h)[mask]), [np.sum((ax3 * h * t[:, :, q, :])[mask]) for q in range(2)]

    h0, t0 = sums()
    m.step(48)
    h1, t1 = sums()
    print(f"  thickness {abs(h1 - h0) / h0:.1e}, tracers " +
          " ".join(f"{abs(b - a) / abs(a):.1e}" for a, b in zip(t0, t1)), flush=True)
    assert abs(h1 - h0) <= 1e-12 * h0
    assert all(abs(b - a) <= 1e-12 * abs(a) for a, b in zip(t0, t1))
    m.close()


def test_remap_negative_thickness_and_refusals():
    print("case remap negative thickness, refusals", flush=True)
    grid = S.Grid(4)
    st = ics.synthetic_state(grid, ntracer=1, lmax=8, seed_index=4)
    h = np.array(st["h"])
    h[5, 5, 3] = -1.0
    m = S.Model(grid, ntracer=1, dt=400.0, grav=S.REF_GRAV, advection="remap")
    m.init_state(h, st["u"], st["tracer"], hs=st["hs"])
    with pytest.raises(S.SwmError, match="remapping"):
        m.step(1)
    with pytest.raises(S.SwmError):
        m.adj_forward(4, 2)
    m.close()
    with pytest.raises(S.SwmError):
        S.Model(grid, ntracer=1, dt=400.0, grav=S.REF_GRAV, advection="remap", precision=S.SWM_MIXED)
    with pytest.raises(S.SwmError):
        S.Model(grid, ntracer=1, dt=400.0, grav=S.REF_GRAV, advection="remap", members=2)
    f = S.Model(grid, ntracer=1, dt=400.0, grav=S.REF_GRAV)
    with pytest.raises(S.SwmError):
        f.set_remap_limiter(False, False)
    f.close()
