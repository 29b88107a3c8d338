"""Quick-plot output packed from the device state (pack_swm_qp,
swm/source/swm_qp_module.F:309-370): test case 5 at level 4 with two tracers,
FP64 and mixed precision.

Every field of a record is the model's FP64 value of the time level rounded
to FP32 (bit for bit against the library's own fields: thickness h, height
h + hs, relative and absolute vorticity, divergence, kinetic energy, tracers
1-2; tracers 3-5 zero as the model carries two), and, for the FP64 path,
within one FP32 unit in the last place (plus 1e-11 of the field's largest
value, the FP64 step tolerance) of the oracle's fields after the same steps.
"""
import numpy as np
import pytest
from scipy.io import netcdf_file

import swm_mi355x as S
from oracle.oracle import Oracle, grid_arrays_from

pytestmark = pytest.mark.gpu

PAIRS = (("thickness", "h"), ("relative", "relative"), ("divergence", "div"), ("absolute", "eta"),
         ("kinetic_energy", "ke"))


def _packed(grid, x):
    m = (grid.array("grid_center_mask") > 0.5).ravel(order="F")
    return np.asarray(x).ravel(order="F")[m]


def _model_fields(m, hs):
    f = {q: m.field(n) for q, n in PAIRS}
    f["height"] = f["thickness"] + hs
    t = m.field("tracer")
    for k in range(t.shape[2]):
        f[f"tracer_{k + 1}"] = t[:, :, k, :]
    return f


@pytest.mark.parametrize("precision", [S.SWM_FP64, S.SWM_MIXED])
def test_qp_pack_from_device(tmp_path, precision):
    print(f"case qp pack g=4 swtc5 precision={precision}", flush=True)
    grid = S.Grid(4)
    st = S.swtc5_state(grid, 2)
    m = S.Model(grid, ntracer=2, dt=100.0, grav=S.REF_GRAV, precision=precision, advection="centered")
    m.init_state(st["h"], st["u"], st["tracer"], hs=st["hs"])
    q = S.QpFile(grid, "SWTC#5", 0, str(tmp_path / "qp.nc"))
    expect = []
    for nsteps in (0, 6):
        if nsteps:
            m.step(nsteps)
        q.pack(m)
        q.write(m.time_seconds)
        expect.append(_model_fields(m, st["hs"]))
    assert q.records == 2
    nc = netcdf_file(str(tmp_path / "qp.nc"), "r", mmap=False)
    v = nc.variables
    assert v["end_period"][1] == 600 / 86400.0
    for r, f in enumerate(expect):
        for name in S.QP_FIELDS:
            got = v[name][r]
            if name in f:
                assert np.array_equal(got, _packed(grid, f[name]).astype(np.float32)), (r, name)
            else:
                assert not got.any(), (r, name)
    if precision == S.SWM_FP64:
        orc = Oracle(grid_arrays_from(grid), 2, 100.0, S.REF_GRAV, hs=st["hs"], scheme="centered")
        orc.init_state(st["h"], st["u"], st["tracer"])
        orc.step(6)
        ref = {q: orc.field(n) for q, n in PAIRS}
        ref["height"] = ref["thickness"] + st["hs"]
        t = orc.field("tracer")
        ref["tracer_1"], ref["tracer_2"] = t[:, :, 0, :], t[:, :, 1, :]
        for name, x in ref.items():
            want = _packed(grid, x).astype(np.float32)
            tol = np.spacing(np.abs(want)) + np.float32(1e-11 * np.abs(x).max())
            assert np.all(np.abs(v[name][1] - want) <= tol), name
    nc.close()
    q.close()
    m.close()
