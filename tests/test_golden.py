"""Host set-up and oracle against the golden vectors of the compiled reference
(tests/golden/, written by tests/golden/make_golden.py).

Bars: bit for bit everywhere -- grid points of our generator (to the
reference generator's 20 printed decimals), metrics, masks and index maps,
halo gather tables, single operators, one multigrid V-cycle, one elliptic
solve and whole-model trajectories over 12 steps, on every slot including the
halos -- because the oracle and the host set-up keep the reference's
floating-point operation order.
"""
import os

import numpy as np
import pytest

import swm_mi355x as S
from oracle import oracle as OR
from oracle.oracle import Oracle, grid_arrays_from, remap_arrays_from

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name))


def grid_for(g):
    if g == 4:
        return S.Grid(4, p_list=gold("setup_g4.npz")["p_list"])
    return S.Grid(g)


def test_generator_matches_reference_generator():
    # the reference generator is fixed at level 5 and prints 20 decimals
    ref = gold("gridgen_g5.npz")["p_list"]
    mine = S.grid_points(5)
    assert mine.shape == ref.shape
    assert np.abs(mine - ref).max() <= 1e-19


@pytest.mark.parametrize("g", [3, 4])
def test_setup_bit_exact(g):
    ref = gold(f"setup_g{g}.npz")
    grid = grid_for(g)
    assert grid.array("tot_area") == ref["tot_area"][0]
    for name in ("p", "c", "area", "area_inv", "corner_area", "corner_area_inv", "sub_area", "dc", "normal",
                 "tangent", "v_weights", "s_weights", "l_weights", "f", "grid_center_mask", "grid_corner_mask",
                 "polygon_type", "grid_center_index", "grid_corner_index", "w", "trans_to_X1", "trans_to_X2"):
        mine = grid.array(name).reshape(-1, order="F")
        assert np.array_equal(mine, ref[name]), name


@pytest.mark.parametrize("g", [1, 2, 3, 4, 5, 6, 7, 8])
def test_halo_tables_bit_exact(g):
    ref = gold(f"wrap_g{g}.npz")
    grid = S.Grid(g)
    for kind in ("face", "vrtx", "edge"):
        d, s = grid.array(kind + "_dst"), grid.array(kind + "_src")
        o = np.argsort(d)
        assert np.array_equal(d[o], ref[kind + "_dst"]), kind
        assert np.array_equal(s[o], ref[kind + "_src"]), kind


def test_halo_sources_are_never_destinations():
    # the device fills halos with an in-place gather; this must be race free
    grid = S.Grid(6)
    for kind in ("face", "vrtx", "edge"):
        d, s = grid.array(kind + "_dst"), grid.array(kind + "_src")
        assert not np.isin(s[s >= 0], d).any()


def _oracle(g, nt=1, dt=900.0):
    grid = grid_for(g)
    return grid, Oracle(grid_arrays_from(grid), nt, dt, S.REF_GRAV)


def test_oracle_operators_bit_exact():
    g = 3
    iim = 2 + 2 ** g
    ns = iim * iim * 10
    raw = gold("ops_g3.npz")["raw"]
    sizes = [("s1", ns), ("s2", ns), ("e1", 3 * ns), ("vc", 4 * ns), ("cc", 2 * ns), ("o1", ns), ("o2", ns),
             ("o3", ns), ("o4", ns), ("o5", ns), ("oc", 2 * ns), ("ov", 4 * ns), ("oe", 3 * ns), ("oe2", 3 * ns),
             ("oe3", 3 * ns), ("oe4", 3 * ns)]
    a, o = {}, 0
    for k, n in sizes:
        a[k] = raw[o:o + n]
        o += n
    assert o == raw.size
    _, orc = _oracle(g)
    c = (iim, iim, 10)
    e = (3,) + c
    checks = [
        ("orc_div_of_an_edge_flux", (a["e1"],), c, "o1"),
        ("orc_laplacian", (a["s1"],), c, "o2"),
        ("orc_curl_of_a_vector", (a["vc"],), c, "o3"),
        ("orc_div_of_a_vector", (a["vc"],), c, "o4"),
        ("orc_center_ke", (a["cc"],), c, "o5"),
        ("orc_average_to_corners", (a["s1"],), (2,) + c, "oc"),
        ("orc_velocity", (a["s1"], a["s2"]), (2, 2) + c, "ov"),
        ("orc_centered_mass_flux", (a["cc"], a["vc"]), e, "oe"),
        ("orc_upstream_mass_flux", (a["vc"], a["s2"]), e, "oe2"),
        ("orc_vorticity_flux", (a["e1"], a["s2"], a["s1"]), e, "oe3"),
        ("orc_divergence_curl", (a["cc"], a["s1"], a["vc"], a["s2"]), e, "oe4"),
    ]
    for fn, args, shape, ref in checks:
        out = orc.op(fn, *args, out_shape=shape)
        assert np.array_equal(out.reshape(-1, order="F"), a[ref]), fn


def _mg_gold():
    g = 4
    iim = 2 + 2 ** g
    n = iim * iim * 2 * 10
    raw = gold("mg_g4.npz")["raw"]
    shp = (iim, iim, 2, 10)
    return [raw[k * n:(k + 1) * n].reshape(shp, order="F") for k in range(4)]


def test_oracle_multigrid_vcycle_bit_exact():
    beta, alpha, _, _ = _mg_gold()
    _, orc = _oracle(4)
    out = orc.multigrid(beta, np.zeros_like(beta))
    assert np.array_equal(out, alpha)


def test_oracle_elliptic_bit_exact():
    beta, _, rhs, xo = _mg_gold()
    _, orc = _oracle(4)
    st, r, x = orc.elliptic(beta.copy(), np.zeros_like(beta))
    assert st == 0
    assert np.array_equal(r, rhs)
    assert np.array_equal(x, xo)


@pytest.mark.parametrize("g", [3, 4])
def test_oracle_trajectory_bit_exact(g):
    ref = gold(f"run_g{g}.npz")
    nt = 5
    dt = float(ref["dt"])
    grid, orc = _oracle(g, nt, dt)
    orc.init_state(ref["ic_h"], ref["ic_u"], ref["ic_tracer"])
    done = 0
    for s in (1, 2, 3, int(ref["nsteps"])):
        orc.step(s - done)
        done = s
        for name in ("h", "eta", "div", "psi", "chi", "ke", "relative", "u", "tracer"):
            mine = orc.field(name).reshape(-1, order="F")
            assert np.array_equal(mine, ref[f"s{s}_{name}"]), f"step {s} {name}"


def test_oracle_compensated_sums_track_the_reference():
    # the oracle's compensated-sum mode (elliptic mean removal and residual
    # norms summed with Neumaier's compensation instead of in memory order)
    # differs from the reference only by the rounding of those sums: over 12
    # steps at g=4 every field within 1e-10 of the reference's (divergence
    # 1.3e-11, the rest below 1e-12), and the same V-cycles as the
    # reference-order oracle
    ref = gold("run_g4.npz")
    grid = grid_for(4)
    dt = float(ref["dt"])
    a = Oracle(grid_arrays_from(grid), 5, dt, S.REF_GRAV, sums="compensated")
    b = Oracle(grid_arrays_from(grid), 5, dt, S.REF_GRAV)
    for o in (a, b):
        o.init_state(ref["ic_h"], ref["ic_u"], ref["ic_tracer"])
    n = int(ref["nsteps"])
    for _ in range(n):
        a.step(1)
        b.step(1)
        assert a.last_vcycles == b.last_vcycles
    for name in ("h", "eta", "div", "psi", "chi", "ke", "u", "tracer"):
        mine = a.field(name).reshape(-1, order="F")
        r = ref[f"s{n}_{name}"]
        assert np.abs(mine - r).max() <= 1e-10 * np.abs(r).max(), name
    assert not np.array_equal(a.field("psi"), b.field("psi"))   # the mode is in effect


def test_oracle_stops_where_the_reference_stops():
    # with this state the reference's elliptic solver fails to reach its
    # criterion in 12 V-cycles and stops the program (elliptic_solver.F:176-178)
    ref = gold("run_g3.npz")
    stop = int(ref["abort_step"])
    assert stop > 0
    _, orc = _oracle(3, 5, float(ref["dt"]))
    orc.init_state(ref["ic_h"], ref["ic_u"], ref["ic_tracer"])
    orc.step(stop - 1)
    with pytest.raises(RuntimeError):
        orc.step(1)


@pytest.mark.parametrize("scheme", ["centered", "upwind"])
def test_oracle_advection_schemes_bit_exact(scheme):
    # step_dt with l_centered / l_upwind (step.F:221-327, :495-602) against the
    # compiled reference, 12 steps at g = 4
    ref = gold(f"run_g4_{scheme}.npz")
    nt, dt = 5, float(ref["dt"])
    grid = grid_for(4)
    orc = Oracle(grid_arrays_from(grid), nt, dt, S.REF_GRAV, scheme=scheme)
    orc.init_state(ref["ic_h"], ref["ic_u"], ref["ic_tracer"])
    done = 0
    for s in (1, 2, 3, int(ref["nsteps"])):
        orc.step(s - done)
        done = s
        for name in ("h", "eta", "div", "psi", "chi", "ke", "relative", "u", "tracer"):
            mine = orc.field(name).reshape(-1, order="F")
            assert np.array_equal(mine, ref[f"s{s}_{name}"]), f"{scheme} step {s} {name}"


# the tracer limiter's switches in the reference's runs: remap_mass_and_tracers
# (remap.F:1072-1604) leaves l_limit_tracers_min / _max without a value when
# the caller does not pass them, and get_remap_flux does not, so each build
# of the reference uses what its stack holds.  The g = 4 build applies the
# two-sided limit (l_limit_tracers = .true.) in every step; the g = 3 build
# in the first step only, and none after.
REMAP_LIMITER = {3: lambda step: (1, 1) if step == 1 else (0, 0), 4: lambda step: (1, 1)}


@pytest.mark.parametrize("g", [3, 4])
def test_oracle_remap_bit_exact(g):
    """step_dt with l_remap (get_remap_flux, step.F:687-771, over
    remap.F:176-1604) against the compiled reference, 12 steps, every field
    bit for bit, with the remapping quantities of the library's host set-up
    (init_remap) and the limiter switches each reference build used."""
    ref = gold(f"run_g{g}_remap.npz")
    nt, dt = 5, float(ref["dt"])
    grid = grid_for(g)
    orc = Oracle(grid_arrays_from(grid), nt, dt, S.REF_GRAV, hs=ref["ic_hs"], scheme="remap",
                 remap=remap_arrays_from(grid), limiter=(1, 1))
    orc.init_state(ref["ic_h"], ref["ic_u"], ref["ic_tracer"])
    done = 0
    for s in (1, 2, 3, int(ref["nsteps"])):
        while done < s:
            done += 1
            OR.lib().orc_set_remap_limiter(orc._h, *REMAP_LIMITER[g](done))
            orc.step(1)
        for name in ("h", "eta", "div", "psi", "chi", "ke", "relative", "u", "tracer"):
            mine = orc.field(name).reshape(-1, order="F")
            assert np.array_equal(mine, ref[f"s{s}_{name}"]), f"g={g} step {s} {name}"


def test_oracle_remap_limiter_pinned():
    """The limiter switches matter: with the two-sided limit in every step the
    g = 3 tracers leave that build's trajectory at step 2 while the thickness
    stays on it (the limiter acts on tracers only)."""
    ref = gold("run_g3_remap.npz")
    grid = grid_for(3)
    orc = Oracle(grid_arrays_from(grid), 5, float(ref["dt"]), S.REF_GRAV, hs=ref["ic_hs"], scheme="remap",
                 remap=remap_arrays_from(grid), limiter=(1, 1))
    orc.init_state(ref["ic_h"], ref["ic_u"], ref["ic_tracer"])
    orc.step(2)
    assert np.array_equal(orc.field("h").reshape(-1, order="F"), ref["s2_h"])
    assert not np.array_equal(orc.field("tracer").reshape(-1, order="F"), ref["s2_tracer"])


def test_oracle_dissipation_bit_exact():
    """init_diss_op and dissipation (swm_dissipation_operator.F:50-443) of a
    pseudo-random corner vector field, norder 1 and 2, against the compiled
    reference on the shipped g = 4 grid, every slot."""
    d = gold("diss_g4.npz")
    grid = S.Grid(4, p_list=d["p_list"])
    n = 4 * grid.iim * grid.jjm * 10
    vin, o1, o2 = [d["raw"][k * n:(k + 1) * n].reshape((2, 2, grid.iim, grid.jjm, 10), order="F") for k in range(3)]
    orc = Oracle(grid_arrays_from(grid), 1, 100.0, S.REF_GRAV)
    orc.set_diss(grid)
    assert np.array_equal(orc.dissipation(1, 2.5e12, vin), o1)
    assert np.array_equal(orc.dissipation(2, 2.5e12, vin), o2)
    orc.close()


def test_oracle_budgets_against_reference():
    """compute_budgets of the reference (budgets.F:31-247) along 12 FCT steps at
    g = 4: its real*4 records (relative 1e-6) and the values it prints with 9
    digits (mass to 1e-8; log10 energy / enstrophy change and tracer variance
    ratios, which the reference derives from its t = 0 values)."""
    ref = gold("budgets_g4.npz")
    rec = ref["records"]
    nt, dt = 5, float(ref["dt"])
    grid = grid_for(4)
    orc = Oracle(grid_arrays_from(grid), nt, dt, S.REF_GRAV)
    orc.init_state(ref["ic_h"], ref["ic_u"], ref["ic_tracer"])
    b0 = None
    for s in range(int(ref["nsteps"]) + 1):
        if s:
            orc.step(1)
        b = orc.budgets()
        mine = [b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7]]
        cols = [0, 1, 2, 3, 4, 7, 10, 11]
        for v, c in zip(mine, cols):
            r = float(rec[s, c])
            assert abs(v - r) <= 1e-6 * abs(r) + 1e-30, (s, c, v, r)
        assert abs(b[0] - ref["mass"][s]) <= 1e-8 * ref["mass"][s]
        if b0 is None:
            b0 = b
            continue
        eps = 1.0e-15
        te_diff = abs(b[4] - b0[4]) / (b0[4] + eps)
        pe_diff = abs(b[5] - b0[5]) / (b0[5] + eps)
        assert abs(np.log10(te_diff + eps) - ref["lte_diff"][s]) <= 1e-6 * abs(ref["lte_diff"][s])
        assert abs(np.log10(pe_diff + eps) - ref["lpenst_diff"][s]) <= 1e-6 * abs(ref["lpenst_diff"][s])
        assert abs(b[8] / b0[8] - ref["mtr1sq_diff"][s]) <= 1e-8
        assert abs(b[9] / b0[9] - ref["mtr2sq_diff"][s]) <= 1e-8
