"""Device operators, halo fill and solver entry points (called through the C
ABI on host arrays) against the golden vectors of the compiled reference and
the oracle.

Bars: bit for bit on every slot for the single operators, the halo fill and
one multigrid V-cycle (no reductions, same operation order); for the elliptic
solve, whose area-weighted mean and residual norms are summed in another
order on the device, a relative difference of 1e-12 and the same V-cycle
count.
"""
import os

import numpy as np
import pytest

import swm_mi355x as S
from oracle.oracle import Oracle, grid_arrays_from

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL_ELLIPTIC = 1e-12


def _ops_gold():
    g = 3
    iim = 2 + 2 ** g
    ns = iim * iim * 10
    raw = np.load(os.path.join(GOLD, "ops_g3.npz"))["raw"]
    sizes = [("s1", 1), ("s2", 1), ("e1", 3), ("vc", 4), ("cc", 2), ("o1", 1), ("o2", 1), ("o3", 1), ("o4", 1),
             ("o5", 1), ("oc", 2), ("ov", 4), ("oe", 3), ("oe2", 3), ("oe3", 3), ("oe4", 3)]
    lead = {1: (), 2: (2,), 3: (3,), 4: (2, 2)}
    a, o = {}, 0
    for k, n in sizes:
        a[k] = raw[o:o + n * ns].reshape(lead[n] + (iim, iim, 10), order="F")
        o += n * ns
    assert o == raw.size
    return a


def _op_cases(grid, a):
    m = {n: grid.array(n) for n in ("area_inv", "v_weights", "l_weights", "corner_area_inv", "s_weights", "sub_area",
                                    "dc", "normal", "tangent")}
    return [
        ("div_of_an_edge_flux", (m["area_inv"], a["e1"]), "o1"),
        ("laplacian_of_a_scalar", (m["l_weights"], a["s1"]), "o2"),
        ("curl_of_a_vector", (m["area_inv"], m["v_weights"], a["vc"]), "o3"),
        ("div_of_a_vector", (m["area_inv"], m["v_weights"], a["vc"]), "o4"),
        ("calculate_center_kinetic_energy", (m["sub_area"], m["area_inv"], a["cc"]), "o5"),
        ("average_to_cell_corners", (m["s_weights"], a["s1"]), "oc"),
        ("calculate_velocity", (m["corner_area_inv"], m["v_weights"], a["s1"], a["s2"]), "ov"),
        ("centered_mass_flux", (m["dc"], m["normal"], a["cc"], a["vc"], a["s2"]), "oe"),
        ("upstream_mass_flux", (m["dc"], m["normal"], a["vc"], a["s2"]), "oe2"),
        ("centered_vorticity_flux", (a["e1"], a["s2"], a["s1"]), "oe3"),
        ("centered_divergence_curl", (m["dc"], m["tangent"], a["cc"], a["s1"], a["vc"], a["s2"]), "oe4"),
    ], m


def test_operators_match_reference_bit_for_bit():
    """All ten panels in one call, and panel by panel as the reference calls
    them (nsdm = 1), against the reference's outputs on every slot."""
    print("case operators g=3 golden", flush=True)
    a = _ops_gold()
    grid = S.Grid(3)
    model = S.Model(grid, ntracer=1, dt=1800.0, grav=S.REF_GRAV)
    cases, _ = _op_cases(grid, a)
    report = {}
    for name, args, ref in cases:
        out = model.op(name, *args)
        bad = out != a[ref]
        if bad.any():
            report[name] = (int(bad.sum()), np.argwhere(bad)[:4].tolist(), float(np.abs(out - a[ref]).max()))
        for k in (0, 9):
            one = model.op(name, *[np.asfortranarray(x[..., k]) for x in args])
            if not np.array_equal(one, a[ref][..., k]):
                report[f"{name} panel {k + 1}"] = "differs"
    assert not report, f"operators differing from the reference (count, first slots, max diff): {report}"


def test_gradient_and_k_cross_v_compose_to_reference_velocity():
    # calculate_velocity = gradient(chi) + k_cross_v(gradient(psi)) (operators.F:1402-1514)
    a = _ops_gold()
    grid = S.Grid(3)
    model = S.Model(grid, ntracer=1, dt=1800.0, grav=S.REF_GRAV)
    _, m = _op_cases(grid, a)
    gchi = model.op("gradient", m["corner_area_inv"], m["v_weights"], a["s2"])
    gpsi = model.op("gradient", m["corner_area_inv"], m["v_weights"], a["s1"])
    kx = model.op("k_cross_v", gpsi)
    assert np.array_equal(kx[0], -gpsi[1]) and np.array_equal(kx[1], gpsi[0])
    assert np.array_equal(gchi + kx, a["ov"])


def test_average_to_cell_corners_other_modes():
    # flag = .false. (fixed weights 0.3333333) and min_value = .true.
    # (operators.F:1052-1106), written out in numpy on the interior; ghost
    # rows and columns carry xin(2,2) of the panel
    a = _ops_gold()
    grid = S.Grid(3)
    model = S.Model(grid, ntracer=1, dt=1800.0, grav=S.REF_GRAV)
    sw = grid.array("s_weights")
    x = a["s1"]
    n = grid.iim
    I = slice(1, n - 1)
    Im = slice(0, n - 2)
    f = 0.3333333
    out = model.op("average_to_cell_corners", sw, x, flag=False)
    assert np.array_equal(out[0, I, I], f * x[I, I] + f * x[Im, I] + f * x[Im, Im])
    assert np.array_equal(out[1, I, I], f * x[I, I] + f * x[Im, Im] + f * x[I, Im])
    mn = model.op("average_to_cell_corners", sw, x, min_value=True)
    assert np.array_equal(mn[0, I, I], np.minimum(np.minimum(sw[0, 0, I, I] * x[I, I], sw[1, 0, I, I] * x[Im, I]),
                                                  sw[2, 0, I, I] * x[Im, Im]))
    assert np.array_equal(mn[1, I, I], np.minimum(np.minimum(sw[0, 1, I, I] * x[I, I], sw[1, 1, I, I] * x[Im, Im]),
                                                  sw[2, 1, I, I] * x[I, Im]))
    for o in (out, mn):
        ghost = np.ones((n, n), bool)
        ghost[I, I] = False
        for k in range(10):
            assert np.all(o[:, ghost, k] == x[1, 1, k])
    with pytest.raises(S.SwmError):
        model.op("average_to_cell_corners", sw, x, min_value=False)


@pytest.mark.parametrize("g", [3, 5])
def test_halo_fill_matches_oracle(g):
    print(f"case wrap g={g}", flush=True)
    grid = S.Grid(g)
    model = S.Model(grid, ntracer=1, dt=900.0, grav=S.REF_GRAV)
    orc = Oracle(grid_arrays_from(grid), 1, 900.0, S.REF_GRAV)
    rng = np.random.default_rng(g)
    iim = grid.iim
    # face, three layers
    x = np.asfortranarray(rng.standard_normal((iim, iim, 3, 10)))
    mine = model.wrap("face", x.copy(order="F"))
    for l in range(3):
        ref = orc.wrap("face", np.asfortranarray(x[:, :, l, :]))
        assert np.array_equal(mine[:, :, l, :], ref), f"face layer {l}"
    # corners with two space components
    v = np.asfortranarray(rng.standard_normal((2, 2, iim, iim, 10)))
    assert np.array_equal(model.wrap("vrtx", v.copy(order="F"), nspace=2), orc.wrap("vrtx", v.copy(order="F"), 2))
    # edges
    e = np.asfortranarray(rng.standard_normal((3, iim, iim, 10)))
    assert np.array_equal(model.wrap("edge", e.copy(order="F")), orc.wrap("edge", e.copy(order="F")))


def _mg_gold():
    g = 4
    iim = 2 + 2 ** g
    n = iim * iim * 2 * 10
    raw = np.load(os.path.join(GOLD, "mg_g4.npz"))["raw"]
    return [raw[k * n:(k + 1) * n].reshape((iim, iim, 2, 10), order="F") for k in range(4)]


def _grid4():
    return S.Grid(4, p_list=np.load(os.path.join(GOLD, "setup_g4.npz"))["p_list"])


def test_vcycle_matches_reference_bit_for_bit():
    print("case multigrid g=4 golden", flush=True)
    beta, alpha, _, _ = _mg_gold()
    model = S.Model(_grid4(), ntracer=1, dt=900.0, grav=S.REF_GRAV)
    out = model.multigrid(beta, np.zeros_like(beta))
    d = np.abs(out - alpha)
    assert np.array_equal(out, alpha), f"{int((d > 0).sum())} slots differ, max {d.max()}"
    # one variable at a time gives the same result
    out0 = model.multigrid(np.asfortranarray(beta[:, :, :1, :]), np.zeros_like(beta[:, :, :1, :]))
    assert np.array_equal(out0[:, :, 0, :], alpha[:, :, 0, :])


def test_elliptic_matches_reference():
    print("case elliptic g=4 golden", flush=True)
    beta, _, rhs_ref, x_ref = _mg_gold()
    grid = _grid4()
    model = S.Model(grid, ntracer=1, dt=900.0, grav=S.REF_GRAV)
    r, x = model.elliptic(beta, np.zeros_like(beta))
    assert np.abs(r - rhs_ref).max() <= TOL_ELLIPTIC * np.abs(rhs_ref).max()
    assert np.abs(x - x_ref).max() <= TOL_ELLIPTIC * np.abs(x_ref).max()
    orc = Oracle(grid_arrays_from(grid), 1, 900.0, S.REF_GRAV)
    st, _, _ = orc.elliptic(beta.copy(), np.zeros_like(beta))
    assert st == 0
    assert model.last_vcycles == orc.last_vcycles


def test_entry_point_errors():
    grid = S.Grid(3)
    model = S.Model(grid, ntracer=1, dt=900.0, grav=S.REF_GRAV)
    with pytest.raises(S.SwmError):
        S._check(S.lib.swm_wrap(model._h, 7, 1, 1, None))
    with pytest.raises(S.SwmError):
        S._check(S.lib.swm_elliptic(model._h, 3, None, None))
    with pytest.raises(S.SwmError):
        S._check(S.lib.swm_op_laplacian_of_a_scalar(model._h, 1, grid.iim, grid.iim, 10, 7, None, None, None))
    lw = grid.array("l_weights")
    x = np.zeros((grid.iim, grid.iim, 10), order="F")
    with pytest.raises(S.SwmError):   # npad other than 1
        S._check(S.lib.swm_op_laplacian_of_a_scalar(model._h, 2, grid.iim, grid.iim, 10, 7, S._ptr(lw), S._ptr(x),
                                                    S._ptr(x)))


def test_operator_bench_leaves_the_state_and_reports_bytes():
    print("case operator microbench", flush=True)
    grid = S.Grid(4)
    m = S.Model(grid, ntracer=1, dt=400.0, grav=S.REF_GRAV)
    m.init_synthetic(lmax=8)
    m.step(2)
    before = {n: m.field(n) for n in ("h", "eta", "div", "u", "psi")}
    per = {"div_of_an_edge_flux": 40.0 * grid.max_ig, "curl_of_a_vector": 144.0 * grid.max_ig,
           "laplacian_of_a_scalar": 72.0 * grid.max_ig,
           "face_halo_fill": 16.0 * len(grid.array("face_dst"))}
    for op in S.Model.OPERATOR_BENCH[:4]:
        us, nbytes = m.operator_bench(op, reps=5)
        assert us > 0.0 and nbytes == per[op], (op, us, nbytes)
    # the remapping and dissipation kernels need their contexts
    for op in ("remap_transport", "remap_fields", "dissipation"):
        with pytest.raises(S.SwmError):
            m.operator_bench(op, reps=5)
    rm = S.Model(grid, ntracer=2, dt=400.0, grav=S.REF_GRAV, advection="remap")
    rm.init_synthetic(lmax=8)
    rm.step(2)
    rm.init_diss_op()
    rb = {n: rm.field(n) for n in ("h", "u", "tracer")}
    per_rm = {"remap_transport": (1128.0 + 48.0 * 2) * grid.max_ig, "remap_fields": (288.0 + 32.0 * 2) * grid.max_ig,
              "dissipation": 368.0 * grid.max_ig}
    for op, nb in per_rm.items():
        us, nbytes = rm.operator_bench(op, reps=5)
        assert us > 0.0 and nbytes == nb, (op, us, nbytes)
    for n, x in rb.items():
        assert np.array_equal(rm.field(n), x), n
    rm.close()
    for n, x in before.items():
        assert np.array_equal(m.field(n), x), n
    mix = S.Model(grid, ntracer=1, dt=400.0, grav=S.REF_GRAV, precision=S.SWM_MIXED)
    mix.init_synthetic(lmax=8)
    with pytest.raises(S.SwmError):
        mix.operator_bench("face_halo_fill")


def test_global_real_sum():
    # global_reductions.F:77-98 sum(X * MASK); compensated, so within a few
    # roundings of the exactly rounded sum (math.fsum) whatever the order
    import math
    print("case global_real_sum", flush=True)
    grid = S.Grid(4)
    m = S.Model(grid, ntracer=1, dt=400.0, grav=S.REF_GRAV)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(200003) * 10.0 ** rng.integers(-8, 8, 200003)
    mask = (rng.random(200003) < 0.7).astype(np.float64)
    exact = math.fsum(x * mask)
    got = m.global_real_sum(x, mask)
    assert abs(got - exact) <= 4e-16 * np.sum(np.abs(x * mask)), (got, exact)
    assert m.global_real_sum(x) == pytest.approx(math.fsum(x), rel=0, abs=4e-16 * np.sum(np.abs(x)))
    # the area of the real cells, as grid_metrics.F sums it
    area = grid.array("area")
    cm = grid.array("grid_center_mask")
    assert abs(m.global_real_sum(area, cm) - grid.array("tot_area")) <= 1e-13 * grid.array("tot_area")
    assert m.global_real_sum(np.zeros(0)) == 0.0
    with pytest.raises(ValueError):
        m.global_real_sum(x, mask[:10])


def test_fct_entry_points_match_oracle():
    # set_mass_fct / set_tracer_fct under their own argument lists, whole
    # sphere, against the oracle's (pinned to the compiled reference through
    # the model trajectories); an AB3 history in permuted slots
    from swm_mi355x import ics
    print("case set_mass_fct / set_tracer_fct", flush=True)
    g, dt = 4, 900.0
    grid = S.Grid(g)
    n = grid.iim
    st = ics.synthetic_state(grid, ntracer=1, lmax=8)
    m = S.Model(grid, ntracer=1, dt=dt, grav=S.REF_GRAV)
    orc = Oracle(grid_arrays_from(grid), 1, dt, S.REF_GRAV)
    rng = np.random.default_rng(5)
    h = st["h"]
    hc = orc.op("orc_average_to_corners", h, out_shape=(2, n, n, 10))
    u = st["u"]
    hist = np.asfortranarray(rng.standard_normal((3, n, n, 3, 10)) * 1e9)
    ti, tw = [2, 3, 1], [23.0 / 12.0, -16.0 / 12.0, 5.0 / 12.0]
    mss = np.zeros((n, n, 1, 2, 10), order="F")
    mss[:, :, 0, 0, :] = h
    msh = np.zeros((3, n, n, 1, 3, 10), order="F")
    msh[:, :, :, 0, :, :] = hist
    args = [grid.array(k) for k in ("area_inv", "area", "dc", "normal")]
    low, flux = m.set_mass_fct(1, 2, ti, tw, *args, mss, hc[:, :, :, None, :], u[:, :, :, :, None, :], msh)
    hist_o = hist.copy(order="F")
    new_o, low_o, flux_o = orc.set_mass_fct(ti, tw, h, hc, u, hist_o)
    assert np.array_equal(low[:, :, 0, :], low_o)
    assert np.array_equal(flux[:, :, :, 0, :], flux_o)
    assert np.array_equal(mss[:, :, 0, 1, :], new_o)
    assert np.array_equal(msh[:, :, :, 0, :, :], hist_o)
    assert np.array_equal(mss[:, :, 0, 0, :], h)   # level n3 untouched
    trc = st["tracer"][:, :, 0, :]
    trh = trc + 0.01 * rng.standard_normal(trc.shape)
    tf, tflux, tml = m.set_tracer_fct(grid.array("area_inv"), grid.array("area"), trc[:, :, None, :],
                                      trh[:, :, None, :], h[:, :, None, :], low, mss[:, :, :, 1, :], flux)
    assert np.array_equal(tf[:, :, 0, :], orc.set_tracer_fct(trc, trh, h, low_o, new_o, flux_o))
    with pytest.raises(S.SwmError):   # the whole sphere only
        S._check(S.lib.swm_op_set_tracer_fct(m._h, n - 1, n, 1, 10, 3, dt, *([None] * 12)))
