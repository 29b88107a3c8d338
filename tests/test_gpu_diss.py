"""The dissipation operator on the device (init_diss_op and dissipation,
swm/source/swm_dissipation_operator.F:50-443): bit for bit against the
compiled reference (tests/golden/diss_g4.npz, norder 1 and 2, every slot)
and against the oracle at level 6 for norder 1, 2 and 3; the calls out of
order are refused."""
import os

import numpy as np
import pytest

import swm_mi355x as S
from oracle.oracle import Oracle, grid_arrays_from

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_dissipation_matches_reference():
    print("case dissipation g=4 vs reference", flush=True)
    d = np.load(os.path.join(GOLD, "diss_g4.npz"))
    grid = S.Grid(4, p_list=d["p_list"])
    n = 4 * grid.iim * grid.jjm * 10
    vin, o1, o2 = [d["raw"][k * n:(k + 1) * n].reshape((2, 2, grid.iim, grid.jjm, 10), order="F") for k in range(3)]
    m = S.Model(grid, ntracer=1, dt=100.0, grav=S.REF_GRAV)
    with pytest.raises(S.SwmError):
        m.dissipation(1, 2.5e12, vin)
    m.init_diss_op()
    assert np.array_equal(m.dissipation(1, 2.5e12, vin), o1)
    assert np.array_equal(m.dissipation(2, 2.5e12, vin), o2)
    with pytest.raises(S.SwmError):
        m.dissipation(0, 2.5e12, vin)
    m.close()


def test_dissipation_matches_oracle_level6():
    print("case dissipation g=6 vs oracle", flush=True)
    grid = S.Grid(6)
    rng = np.random.default_rng(11)
    x = 20.0 * rng.standard_normal((2, 2, grid.iim, grid.jjm, 10))
    m = S.Model(grid, ntracer=1, dt=100.0, grav=S.REF_GRAV)
    m.init_diss_op()
    orc = Oracle(grid_arrays_from(grid), 1, 100.0, S.REF_GRAV)
    orc.set_diss(grid)
    for norder in (1, 2, 3):
        assert np.array_equal(m.dissipation(norder, 1.0e15, x), orc.dissipation(norder, 1.0e15, x)), norder
    m.close()
    orc.close()
