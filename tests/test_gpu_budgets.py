"""Device budgets (swm_budgets) against the reference's compute_budgets output
(budgets.F:31-247, golden records of the compiled reference) and the oracle.

Bars: against the oracle, relative 1e-12 (the device sums in another,
compensated order); against the reference's real*4 records, relative 1e-6.
The global means of absolute vorticity and divergence vanish analytically and
are sums of rounding errors (about 1e-16 of the terms' size), so those two
are held to 1e-12 of the mean magnitude of their field instead.  The mixed
path against the FP64 path after 12 steps: mass 1e-9, energies and
enstrophy 1e-5 (the two forms' elliptic solves stop at other residuals
within the criterion)."""
import os

import numpy as np
import pytest

import swm_mi355x as S
from oracle.oracle import Oracle, grid_arrays_from

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _grid4():
    return S.Grid(4, p_list=np.load(os.path.join(GOLD, "setup_g4.npz"))["p_list"])


def test_budgets_match_reference_and_oracle():
    print("case budgets g=4 fct 12 steps", flush=True)
    ref = np.load(os.path.join(GOLD, "budgets_g4.npz"))
    rec = ref["records"]
    dt, nt = float(ref["dt"]), 5
    grid = _grid4()
    model = S.Model(grid, ntracer=nt, dt=dt, grav=S.REF_GRAV)
    model.init_state(ref["ic_h"], ref["ic_u"], ref["ic_tracer"], hs=ref["ic_hs"])
    orc = Oracle(grid_arrays_from(grid), nt, dt, S.REF_GRAV, hs=ref["ic_hs"])
    orc.init_state(ref["ic_h"], ref["ic_u"], ref["ic_tracer"])
    mask = grid.array("grid_center_mask") == 1.0
    cols = [0, 1, 2, 3, 4, 7, 10, 11]
    for s in range(int(ref["nsteps"]) + 1):
        if s:
            model.step(1)
            orc.step(1)
        b, o = model.budgets(), orc.budgets()
        # scale of the vanishing means (q = 6 eta, q = 7 div)
        scale = {6: np.abs(orc.field("eta")[mask]).mean(), 7: np.abs(orc.field("div")[mask]).mean()}
        for q in range(len(o)):
            tol = 1e-12 * (scale[q] if q in scale else abs(o[q]))
            assert abs(b[q] - o[q]) <= tol + 1e-300, (s, q, b[q], o[q])
        for q, c in enumerate(cols):
            r = float(rec[s, c])
            tol = 1e-12 * scale[q] if q in scale else 1e-6 * abs(r)
            assert abs(b[q] - r) <= tol + 1e-300, (s, c, b[q], r)


def test_mixed_budgets_close_to_fp64():
    ref = np.load(os.path.join(GOLD, "budgets_g4.npz"))
    grid = _grid4()
    out = {}
    for prec in (S.SWM_FP64, S.SWM_MIXED):
        m = S.Model(grid, ntracer=5, dt=400.0, grav=S.REF_GRAV, precision=prec)
        m.init_state(ref["ic_h"], ref["ic_u"], ref["ic_tracer"], hs=ref["ic_hs"])
        m.step(12)
        out[prec] = m.budgets()
    a, b = out[S.SWM_FP64], out[S.SWM_MIXED]
    # mass to round-off; the energies and enstrophy within 1e-5: the two
    # forms' elliptic solves stop at other residuals within the criterion
    assert abs(a[0] - b[0]) <= 1e-9 * abs(a[0]), (a[0], b[0])
    for q in (1, 2, 4, 5, 8, 9):
        assert abs(a[q] - b[q]) <= 1e-5 * abs(a[q]), (q, a[q], b[q])
