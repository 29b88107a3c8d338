"""The adjoint and ensemble forms at the benchmark workloads' sizes.

Adjoint: icosahedral level 7 (163842 cells), FP64, flux-corrected transport
linearised with the limiter's coefficients frozen at the base trajectory,
dt = 50 s (the reference's own 200 s at 10242 cells scaled with the grid
spacing), a 6 h window (432 steps), snapshots every 16 steps with
recomputation.  Objective J = sum over a spherical cap of radius 1500 km of
area * (h - 5000)^2 after the window; control: the initial relative
vorticity, divergence and thickness of the seeded synthetic state.  The
linearised model is that of the forward as it runs (swm_set_linearisation
exact): each step's solve is its V-cycles from the previous solution.
  - the production forward (the reference's criterion, 2-3 V-cycles per
    solve): dot-product test <TL x, y> = <x, AD y> to 1e-10; central
    differences at eps = 1e-6 relative to 2e-6, the FP64 rounding noise of J
    divided by 2 eps (larger eps changes the criterion's cycle counts
    between the two runs);
  - the forward with exactly 3 V-cycles per solve (a smooth function of the
    control): fourth-order central differences at eps = 1e-4 along 8
    seeded directions to 1e-7;
  - the limiter's activity over the window is recorded (the frozen
    linearisation is the derivative where no coefficient is below 1).

Ensemble: 96 members at level 6 (dt 100 s) and 32 at level 7 (dt 50 s, two
substeps), the mixed step with one fixed correction cycle per solve, in
captured graphs with the stream overlap's default start at level g-2 (level 4
for the level-6 group), one stream or one per group:
  - graph replays equal eager stepping of twin groups bit for bit;
  - a member equals a single-member context from the same state bit for bit.
"""
import numpy as np
import pytest

import swm_mi355x as S
from swm_mi355x import adjoint as A

pytestmark = pytest.mark.gpu

RADIUS, GRAV, OMEGA = 6.371e6, 9.81, 7.292e-5


@pytest.fixture(scope="module")
def window7():
    g = 7
    dt = 200.0 * 2.0 ** (5 - g)
    nsteps = int(round(6 * 3600.0 / dt))
    grid = S.Grid(g, radius=RADIUS, omega=OMEGA)

    def model(fixed=0):
        m = S.Model(grid, ntracer=1, dt=dt, grav=GRAV, advection="fct")
        if fixed:
            m.set_fixed_cycles(fixed, 0)
        return m

    syn = S.Model(grid, ntracer=1, dt=dt, grav=GRAV)
    syn.init_synthetic(seed_index=3, lmax=40)
    control = [np.asfortranarray(syn.field(n)) for n in ("relative", "div", "h")]
    m = model()
    m.init_vorticity(*control)
    m.adj_forward(nsteps, 16)
    return grid, m, model, control, nsteps


def _cap(grid):
    return A.cap_mask(grid, np.pi / 6, -np.pi / 2, 1500e3, RADIUS)


@pytest.mark.timeout(300)
def test_adjoint_g7_dot_product(window7):
    grid, m, _, control, nsteps = window7
    print(f"case adjoint g=7 window {nsteps} steps, production forward: dot product", flush=True)
    rng = np.random.default_rng(5)
    mask = grid.array("grid_center_mask") == 1.0
    s = (grid.iim, grid.jjm, 10)
    dc = [np.where(mask, np.sqrt(np.mean(c[mask] ** 2)) * rng.standard_normal(s), 0.0) for c in control]
    tl = m.tl_run(control=dc)
    y = [rng.standard_normal(s), 1e6 * rng.standard_normal(s), 1e7 * rng.standard_normal(s),
         10 * rng.standard_normal((2, 2) + s)]
    a = sum(float(np.sum(u * v)) for u, v in zip(tl, y))
    ad = m.adj_run(*y, control=True)
    b = sum(float(np.sum(u * v)) for u, v in zip(dc, ad))
    below, edges = m.limiter_stats()
    print(f"  <TL x, y> = {a:.15e}, <x, AD y> = {b:.15e}, rel {abs(a - b) / abs(a):.2e}; "
          f"limiter coefficients below 1: {below} of {edges} edge-steps", flush=True)
    assert abs(a - b) <= 1e-10 * abs(a)
    assert edges == 3 * (grid.iim - 2) * (grid.jjm - 2) * 10 * nsteps


@pytest.mark.timeout(300)
def test_adjoint_g7_production_forward_central_differences(window7):
    grid, m, model, control, nsteps = window7
    print(f"case adjoint g=7 window {nsteps} steps, production forward: central differences", flush=True)
    area = grid.array("area")
    mask = grid.array("grid_center_mask") == 1.0
    cap = _cap(grid)
    hN = m.field("h")
    grad = m.adj_run(np.where(cap, 2.0 * area * (hN - 5000.0), 0.0), control=True)
    below, edges = m.limiter_stats()
    assert below == 0, (below, edges)   # the frozen linearisation is the derivative on this window
    obj = lambda h: A.cap_objective(h, area, cap, 5000.0)
    fd = A.finite_difference_check(model, control, grad, obj, nsteps, 8, 8, mask, eps=1e-6)
    print("  eps 1e-6: " + " ".join(f"{r[2]:.1e}" for r in fd), flush=True)
    assert all(r[2] <= 2e-6 for r in fd), fd


@pytest.mark.timeout(600)
def test_adjoint_g7_fixed_cycles_fourth_order_differences(window7):
    grid, _, model, control, nsteps = window7
    print(f"case adjoint g=7 window {nsteps} steps, 3 V-cycles per solve: fourth-order differences", flush=True)
    area = grid.array("area")
    mask = grid.array("grid_center_mask") == 1.0
    cap = _cap(grid)
    m = model(fixed=3)
    m.init_vorticity(*control)
    m.adj_forward(nsteps, 16)
    hN = m.field("h")
    grad = m.adj_run(np.where(cap, 2.0 * area * (hN - 5000.0), 0.0), control=True)
    obj = lambda h: A.cap_objective(h, area, cap, 5000.0)
    fd = A.finite_difference_check(lambda: model(fixed=3), control, grad, obj, nsteps, 8, 8, mask, eps=1e-4,
                                   order=4)
    print("  eps 1e-4, fourth order: " + " ".join(f"{r[2]:.1e}" for r in fd), flush=True)
    assert all(r[2] <= 1e-7 for r in fd), fd


def _group(grid, n, seed0, fixed=1):
    m = S.Model(grid, ntracer=1, dt=200.0 * 2.0 ** (5 - grid.g), grav=GRAV, precision=S.SWM_MIXED, members=n)
    m.set_fixed_cycles(fixed, 8)
    for k in range(n):
        m.select(k)
        m.init_synthetic_member(seed0 + k, 0.01, seed_index=4, lmax=40)
    return m


FIELDS = ("h", "eta", "div", "psi", "chi", "ke", "u", "tracer")


def _member(m, k):
    m.select(k)
    return {n: m.field(n) for n in FIELDS}


@pytest.fixture(scope="module")
def grids67():
    return S.Grid(6, radius=RADIUS, omega=OMEGA), S.Grid(7, radius=RADIUS, omega=OMEGA)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("concurrent", [0, 1])
def test_ensemble_96_g6_32_g7_replay_matches_eager(grids67, concurrent):
    g6, g7 = grids67
    print(f"case ensemble 96 x g6 + 32 x g7, concurrent={concurrent}", flush=True)
    a1, b1 = _group(g6, 96, 0), _group(g7, 32, 96)
    a2, b2 = _group(g6, 96, 0), _group(g7, 32, 96)
    e = S.Ensemble([a2, b2], [1, 2])
    e.set_choice(concurrent=concurrent, steps_per_graph=3, blk_step=256, blk_mg=64)
    # 12 ensemble steps capture the start-up and periodic phases, the rest
    # replay the graphs
    check = {6: [0, 50, 95], 7: [0, 17, 31]}
    for n in (12, 6, 3):
        e.step(n)
        a1.step(n)
        b1.step(2 * n)
        for (m1, m2, g) in ((a1, a2, 6), (b1, b2, 7)):
            for k in check[g]:
                x, y = _member(m1, k), _member(m2, k)
                assert all(np.array_equal(x[f], y[f]) for f in FIELDS), (g, k, n)
    assert a2.step_count == a1.step_count == 21 and b2.step_count == b1.step_count == 42
    # every solve of the timed schedule met the criterion
    for m in (a2, b2):
        for k in range(m.members):
            m.select(k)
            r0, r1, above = m.solver_track()
            assert above == 0 and r0 <= 1e-20 and r1 <= 1e-20, (m.grid.g, k, r0, r1)
    # a member equals a single-member context stepped from the same state
    for (m, g, grid, seed0, k, sub) in ((a2, 6, g6, 0, 50, 1), (b2, 7, g7, 96, 17, 2)):
        one = _group(grid, 1, seed0 + k)
        one.step(21 * sub)
        x, y = _member(one, 0), _member(m, k)
        assert all(np.array_equal(x[f], y[f]) for f in FIELDS), (g, k)
        one.close()
    e.close()
    for m in (a1, b1, a2, b2):
        m.close()
