"""The Fortran side of the boundary: swm_fortran_check drives the device model
through the ISO_C_BINDING modules (fortran/swm_mi355x.F90) the way the
reference driver does (initial wrap_prognostics, then step_dt +
wrap_prognostics per step) and calls operators panel by panel under the
reference's names and argument lists.  Its results must equal, bit for bit,
those of the same library driven from Python, and the operators those of the
oracle.
"""
import os
import subprocess

import numpy as np
import pytest

import swm_mi355x as S
from swm_mi355x import ics
from oracle.oracle import Oracle, grid_arrays_from

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "shallow-water-model_amd", "lib", "swm_fortran_check")


def test_fortran_driver_matches_library(tmp_path):
    print("case fortran g=4 nsteps=3", flush=True)
    g, nt, nsteps, dt = 4, 2, 3, 900.0
    grid = S.Grid(g)
    st = ics.synthetic_state(grid, ntracer=nt, lmax=8, seed_index=4)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(np.array([g, nt, nsteps], dtype="<i4").tobytes())
        f.write(np.array([dt, S.REF_GRAV, S.REF_RADIUS, S.REF_OMEGA], dtype="<f8").tobytes())
        for k in ("h", "hs", "u", "tracer"):
            f.write(np.asfortranarray(st[k], dtype="<f8").tobytes(order="F"))
    (tmp_path / "restarts").mkdir()
    (tmp_path / "qp_output").mkdir()
    r = subprocess.run([EXE, str(fin), str(fout), "restart"], capture_output=True, text=True, timeout=120,
                       cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    n = grid.iim
    raw = np.fromfile(fout, dtype="<f8")
    shapes = [("h", (n, n, 10)), ("eta", (n, n, 10)), ("div", (n, n, 10)), ("psi", (n, n, 10)),
              ("chi", (n, n, 10)), ("u", (2, 2, n, n, 10)), ("tracer", (n, n, nt, 10)), ("lap", (n, n, 10)),
              ("vf", (3, n, n, 10))]
    out, o = {}, 0
    for name, shp in shapes:
        m = int(np.prod(shp))
        out[name] = raw[o:o + m].reshape(shp, order="F")
        o += m
    assert o == raw.size

    model = S.Model(grid, ntracer=nt, dt=dt, grav=S.REF_GRAV)
    model.init_state(st["h"], st["u"], st["tracer"], hs=st["hs"])
    model.step(nsteps)
    for name in ("h", "eta", "div", "psi", "chi", "u", "tracer"):
        assert np.array_equal(out[name], model.field(name)), name

    # write_swm_restart through the Fortran module: the reference's name for
    # the model time, the state of the real cells
    name = S.restart_name(int(nsteps * dt))
    ts, rh, ru, rt = S.read_restart_file(grid, str(tmp_path / name), nt)
    assert ts == int(nsteps * dt)
    m = grid.array("grid_center_mask") == 1.0
    assert np.array_equal(rh[m], out["h"][m])
    assert np.array_equal(ru[:, :, m], out["u"][:, :, m])
    assert np.array_equal(rt.transpose(0, 1, 3, 2)[m], out["tracer"].transpose(0, 1, 3, 2)[m])

    # initialize_swm_qp / pack_swm_qp / write_swm_qp through the Fortran
    # module: the same file as the library makes driven from Python (the
    # creation time in the history attribute aside)
    mine = tmp_path / "mine_qp.nc"
    m0 = S.Model(grid, ntracer=nt, dt=dt, grav=S.REF_GRAV)
    m0.init_state(st["h"], st["u"], st["tracer"], hs=st["hs"])
    q = S.QpFile(grid, "SWTC#5", 0, str(mine))
    q.pack(m0)
    q.write(0)
    m0.step(nsteps)
    q.pack(m0)
    q.write(int(nsteps * dt))
    q.close()
    a, b = (tmp_path / S.qp_name(0)).read_bytes(), mine.read_bytes()
    k = b.index(b"Created: ")
    assert a[:k] == b[:k] and a[k + 39:] == b[k + 39:]

    orc = Oracle(grid_arrays_from(grid), nt, dt, S.REF_GRAV)
    h = model.field("h")
    assert np.array_equal(out["lap"], orc.op("orc_laplacian", h, out_shape=(n, n, 10)))
    vf = orc.op("orc_vorticity_flux", np.ones((3, n, n, 10)), h, model.field("eta"), out_shape=(3, n, n, 10))
    assert np.array_equal(out["vf"], vf)


def test_fortran_drop_ins_match_library_and_oracle(tmp_path):
    """Every drop-in through the Fortran modules (swm_fortran_check ops): the
    13 operators panel by panel, wrap with its keyword arguments, elliptic,
    multigrid, set_mass_fct, set_tracer_fct, global_real_sum.  Operators,
    halo fills, the V-cycle and both FCTs equal the oracle bit for bit
    (operators the oracle has; k_cross_v and gradient equal the library's
    Python path, whose composition is the reference's velocity,
    test_gpu_ops.py); the elliptic solve is the library's bit for bit and the
    oracle's to 1e-12; the global sum is the total area."""
    import math
    print("case fortran drop-ins g=4", flush=True)
    g, dt = 4, 900.0
    grid = S.Grid(g)
    n = grid.iim
    st = ics.synthetic_state(grid, ntracer=1, lmax=8, seed_index=4)
    rng = np.random.default_rng(17)
    s1, s2 = rng.standard_normal((n, n, 10)), 5000.0 + rng.standard_normal((n, n, 10))
    e1 = rng.standard_normal((3, n, n, 10))
    vc = rng.standard_normal((2, 2, n, n, 10))
    cc = 5000.0 + rng.standard_normal((2, n, n, 10))
    hist = rng.standard_normal((3, n, n, 1, 3, 10)) * 1e9
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(np.array([g], dtype="<i4").tobytes())
        f.write(np.array([dt, S.REF_GRAV, S.REF_RADIUS, S.REF_OMEGA], dtype="<f8").tobytes())
        for a in (st["h"], st["hs"], st["u"], st["tracer"], s1, s2, e1, vc, cc, hist):
            f.write(np.asfortranarray(a, dtype="<f8").tobytes(order="F"))
    r = subprocess.run([EXE, "ops", str(fin), str(fout)], capture_output=True, text=True, timeout=120,
                       cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    c3 = (n, n, 10)
    shapes = [("curl", c3), ("divv", c3), ("dive", c3), ("lap", c3), ("kxv", (2, 2) + c3), ("grad", (2, 2) + c3),
              ("avg", (2,) + c3), ("ke", c3), ("vel", (2, 2) + c3), ("cmf", (3,) + c3), ("umf", (3,) + c3),
              ("cvf", (3,) + c3), ("cdc", (3,) + c3),
              ("fw", (n, n, 2, 10)), ("f1l", c3), ("vx", (2, 2, n, n, 1, 10)), ("v1l", (2, 2) + c3),
              ("vs", (2, n, n, 1, 10)), ("e1l", (2, 3) + c3), ("es", (3, n, n, 1, 10)),
              ("rhs", (n, n, 2, 10)), ("xo", (n, n, 2, 10)), ("alpha", (n, n, 2, 10)),
              ("mss", (n, n, 1, 2, 10)), ("mlow", (n, n, 1, 10)), ("mflux", (3, n, n, 1, 10)),
              ("hist", (3, n, n, 1, 3, 10)), ("tfct", (n, n, 1, 10)), ("tflux", (3, n, n, 1, 10)),
              ("tml", (n, n, 1, 10)), ("total", (1,))]
    raw = np.fromfile(fout, dtype="<f8")
    out, o = {}, 0
    for name, shp in shapes:
        k = int(np.prod(shp))
        out[name] = raw[o:o + k].reshape(shp, order="F")
        o += k
    assert o == raw.size
    orc = Oracle(grid_arrays_from(grid), 1, dt, S.REF_GRAV)
    op = lambda name, *a, shape: orc.op(name, *a, out_shape=shape)
    eq = lambda name, ref: np.testing.assert_array_equal(out[name], ref, err_msg=name)
    eq("curl", op("orc_curl_of_a_vector", vc, shape=c3))
    eq("divv", op("orc_div_of_a_vector", vc, shape=c3))
    eq("dive", op("orc_div_of_an_edge_flux", e1, shape=c3))
    eq("lap", op("orc_laplacian", s1, shape=c3))
    eq("avg", op("orc_average_to_corners", s1, shape=(2,) + c3))
    eq("ke", op("orc_center_ke", cc, shape=c3))
    eq("vel", op("orc_velocity", s1, s2, shape=(2, 2) + c3))
    eq("cmf", op("orc_centered_mass_flux", cc, vc, shape=(3,) + c3))
    eq("umf", op("orc_upstream_mass_flux", vc, s2, shape=(3,) + c3))
    eq("cvf", op("orc_vorticity_flux", e1, s2, s1, shape=(3,) + c3))
    eq("cdc", op("orc_divergence_curl", cc, s1, vc, s2, shape=(3,) + c3))
    m = S.Model(grid, ntracer=1, dt=dt, grav=S.REF_GRAV)
    m.init_state(st["h"], st["u"], st["tracer"], hs=st["hs"])
    eq("kxv", m.op("k_cross_v", vc))
    eq("grad", m.op("gradient", grid.array("corner_area_inv"), grid.array("v_weights"), s1))
    # halo fills: every layer and component as the oracle's wrap does it
    w = lambda kind, x, nspace=1: orc.wrap(kind, np.asfortranarray(x), nspace)
    eq("fw", np.stack([w("face", s1), w("face", s2)], axis=2))
    eq("f1l", w("face", s2))
    eq("vx", w("vrtx", vc, 2)[:, :, :, :, None, :])
    eq("v1l", w("vrtx", vc, 2))
    eq("vs", w("vrtx", cc[None], 1)[0][:, :, :, None, :])
    eq("e1l", np.stack([w("edge", e1), w("edge", 2.0 * e1)]))
    es = w("edge", e1)
    es[0, n - 1, n - 2, :] = 0.0   # l_HA_stencil (wrap_data.F:1956-1971)
    es[2, n - 2, n - 1, :] = 0.0
    eq("es", es[:, :, :, None, :])
    # elliptic: the library's bit for bit, the oracle's to 1e-12
    rhs0 = np.stack([1e-5 * s1, 1e-6 * s2], axis=2)
    r_lib, x_lib = m.elliptic(rhs0, np.zeros_like(rhs0))
    eq("rhs", r_lib)
    eq("xo", x_lib)
    _, r_o, x_o = orc.elliptic(rhs0, np.zeros_like(rhs0))
    # the solution is defined up to a constant: the mean removal of the
    # right-hand side (whose mean is large here) rounds differently on the two
    # sides, and the solves carry that difference as a constant
    mask = grid.array("grid_center_mask") == 1.0
    area = grid.array("area")
    for v in range(2):
        a, b = out["xo"][:, :, v, :], x_o[:, :, v, :]
        a = a - np.sum((a * area)[mask]) / np.sum(area[mask])
        b = b - np.sum((b * area)[mask]) / np.sum(area[mask])
        assert np.abs(a - b)[mask].max() <= 1e-12 * np.abs(b)[mask].max(), v
    eq("alpha", orc.multigrid(out["rhs"], np.zeros_like(rhs0)))
    # the two FCT routines
    hc = op("orc_average_to_corners", st["h"], shape=(2,) + c3)
    hist_o = hist[:, :, :, 0, :, :].copy(order="F")
    new_o, low_o, flux_o = orc.set_mass_fct([2, 3, 1], [23.0 / 12.0, -16.0 / 12.0, 5.0 / 12.0], st["h"], hc,
                                            st["u"], hist_o)
    eq("mss", np.stack([st["h"], new_o], axis=2)[:, :, None])
    eq("mlow", low_o[:, :, None])
    eq("mflux", flux_o[:, :, :, None])
    eq("hist", hist_o[:, :, :, None])
    trc = st["tracer"][:, :, 0, :]
    eq("tfct", orc.set_tracer_fct(trc, trc + 0.01 * s1, st["h"], low_o, new_o, flux_o)[:, :, None])
    tf, tflux, tml = m.set_tracer_fct(grid.array("area_inv"), grid.array("area"), trc[:, :, None],
                                      (trc + 0.01 * s1)[:, :, None], st["h"][:, :, None], low_o[:, :, None],
                                      new_o[:, :, None], flux_o[:, :, :, None])
    eq("tflux", tflux)
    eq("tml", tml)
    total = grid.array("tot_area")
    assert abs(out["total"][0] - total) <= 1e-13 * total
    assert abs(out["total"][0] - math.fsum(grid.array("area")[mask])) <= 1e-15 * total


def test_fortran_remap_and_dissipation_match_library(tmp_path):
    """remap_mass_and_tracers (remap_mi355x), dissipation
    (swm_dissipation_operator_mi355x) and the l_remap model through the
    Fortran modules (swm_fortran_check remap): equal, bit for bit, to the
    same library driven from Python."""
    print("case fortran remap g=4", flush=True)
    g, nt, nsteps, dt = 4, 2, 3, 400.0
    grid = S.Grid(g)
    st = ics.synthetic_state(grid, ntracer=nt, lmax=8, seed_index=4)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(np.array([g, nt, nsteps], dtype="<i4").tobytes())
        f.write(np.array([dt, S.REF_GRAV, S.REF_RADIUS, S.REF_OMEGA], dtype="<f8").tobytes())
        for k in ("h", "hs", "u", "tracer"):
            f.write(np.asfortranarray(st[k], dtype="<f8").tobytes(order="F"))
    r = subprocess.run([EXE, "remap", str(fin), str(fout)], capture_output=True, text=True, timeout=120,
                       cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    n = grid.iim
    raw = np.fromfile(fout, dtype="<f8")
    shapes = [("mass", (n, n, 10)), ("trace", (n, n, nt, 10)), ("ef", (3, n, n, 10)), ("mass2", (n, n, 10)),
              ("trace2", (n, n, nt, 10)), ("visc", (2, 2, n, n, 10)), ("h", (n, n, 10)), ("tracer", (n, n, nt, 10))]
    out, o = {}, 0
    for name, shp in shapes:
        k = int(np.prod(shp))
        out[name] = raw[o:o + k].reshape(shp, order="F")
        o += k
    assert o == raw.size
    m = S.Model(grid, ntracer=nt, dt=dt, grav=S.REF_GRAV, advection="remap")
    m.init_state(st["h"], st["u"], st["tracer"], hs=st["hs"])
    h0, u0, t0 = m.field("h"), m.field("u"), m.field("tracer")   # halos filled
    a = m.remap_mass_and_tracers(h0, t0, dt, u0)
    assert np.array_equal(out["mass"], a[0]) and np.array_equal(out["trace"], a[1]) and np.array_equal(out["ef"], a[2])
    m.set_remap_limiter(False, False)
    b = m.remap_mass_and_tracers(h0, t0, dt, u0)
    m.set_remap_limiter(True, True)
    bh, bt = b[0].copy(order="F"), b[1].copy(order="F")
    m.wrap("face", bh)
    for q in range(nt):
        layer = np.asfortranarray(bt[:, :, q, :])
        m.wrap("face", layer)
        bt[:, :, q, :] = layer
    b = m.remap_mass_and_tracers(bh, bt, dt, u0)
    assert np.array_equal(out["mass2"], b[0]) and np.array_equal(out["trace2"], b[1])
    m.init_diss_op()
    assert np.array_equal(out["visc"], m.dissipation(2, 2.5e12, u0))
    m.step(nsteps)
    assert np.array_equal(out["h"], m.field("h")) and np.array_equal(out["tracer"], m.field("tracer"))
    m.close()
