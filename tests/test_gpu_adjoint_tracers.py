"""The tracers in the tangent-linear (TL) and adjoint (AD) model
(swm_tl_run_tracer / swm_adj_run_tracer), on the device, for each advection
scheme: tracer FCT (tracer_fct.F) with its limiter's coefficients frozen at
the base trajectory, and the centred and upstream tracer fluxes with
step_tracer (step.F:278-327).

Tracers do not feed back on the dynamics but depend on it through the mass
flux and the thickness, so the checks cover both directions:
  - dot-product tests <TL x, y> = <x, AD y> with tracers in x and y, over
    every slot of every field, for the window alone (mode 0) and through
    the initial state map (mode 1, initial tracers in the control), to
    1e-10 relative; the tracers' part alone and the dynamics-to-tracer part
    alone as well;
  - the dynamics are unchanged by carrying the tracers: the TL of (h, eta,
    div, u) equals the run without tracers bit for bit, the AD with a zero
    tracer adjoint to rounding (its atomic adds vary in order), and a
    tracer perturbation leaves the dynamics' TL at zero;
  - the gradient of the tracer mass in a spherical cap after the window,
    J = sum_cap area h t, with respect to initial vorticity, divergence,
    thickness and tracers, against central finite differences of the
    nonlinear model (eps = 1e-6 relative) to 1e-7.  For FCT the frozen
    linearisation is the derivative where the tracer limiter is inactive;
    the smooth tracer here has its extrema far from the cap, and the test
    records how many frozen coefficients are below 1.
"""
import numpy as np
import pytest

import swm_mi355x as S

pytestmark = pytest.mark.gpu

G, NSTEPS, EVERY, NT = 4, 8, 3, 2
DT = 200.0 * 2 ** (5 - G)
SCHEMES = ("centered", "fct", "upwind")


def _control(grid):
    p = grid.array("p")
    x, y, z = p[0], p[1], p[2]
    zeta = 1e-5 * (z + 0.5 * x * y + 0.3 * y * y)
    delta = 1e-6 * (x * z - 0.2 * y)
    h = 5000.0 + 100.0 * (x + 0.3 * z * z)
    tr = np.zeros((grid.iim, grid.jjm, NT, 10), order="F")
    tr[:, :, 0, :] = 1.0 + 0.2 * z + 0.1 * x
    tr[:, :, 1, :] = 0.5 + 0.3 * y - 0.1 * z
    return [np.asfortranarray(a) for a in (zeta, delta, h)], tr


def _model(grid, adv):
    m = S.Model(grid, ntracer=NT, dt=DT, grav=S.REF_GRAV, advection=adv)
    m.set_solver(tol=1e-13, relative=True, max_cycles=60)
    return m


@pytest.fixture(scope="module", params=SCHEMES)
def window(request):
    grid = S.Grid(G)
    m = _model(grid, request.param)
    c, tr = _control(grid)
    m.init_vorticity(*c, tracer=tr)
    m.adj_forward(NSTEPS, EVERY)
    m.scheme = request.param
    return grid, m, c, tr


def _dot(a, b):
    return sum(float(np.sum(x * y)) for x, y in zip(a, b))


def _vectors(grid, rng):
    s = (grid.iim, grid.jjm, 10)
    ts = (grid.iim, grid.jjm, NT, 10)
    x = [rng.standard_normal(s), 1e-6 * rng.standard_normal(s), 1e-7 * rng.standard_normal(s),
         0.1 * rng.standard_normal((2, 2) + s), 0.01 * rng.standard_normal(ts)]
    y = [rng.standard_normal(s), 1e6 * rng.standard_normal(s), 1e7 * rng.standard_normal(s),
         10 * rng.standard_normal((2, 2) + s), 100 * rng.standard_normal(ts)]
    return x, y


def test_dot_product_with_tracers(window):
    grid, m, _, _ = window
    print(f"case tracer TL/AD dot product, window, {m.scheme}", flush=True)
    rng = np.random.default_rng(21)
    x, y = _vectors(grid, rng)
    tx = m.tl_run(*x[:4], dtracer=x[4])
    ay = m.adj_run(*y[:4], ltracer=y[4])
    a, b = _dot(tx, y), _dot(x, ay)
    print(f"  <TL x, y> = {a:.15e}, <x, AD y> = {b:.15e}, rel {abs(a - b) / abs(a):.1e}", flush=True)
    assert abs(a - b) <= 1e-10 * abs(a), (a, b)
    # the tracers' own part, and the dynamics' part of the tracers
    z = [np.zeros_like(v) for v in x]
    t_only = m.tl_run(*z[:4], dtracer=x[4])
    a = float(np.sum(t_only[4] * y[4]))
    b = float(np.sum(x[4] * m.adj_run(*z[:4], ltracer=y[4])[4]))
    assert abs(a - b) <= 1e-10 * abs(a), ("tracer -> tracer", a, b)
    assert all(not np.any(t_only[j]) for j in range(4))   # tracers do not feed back
    d_only = m.tl_run(*x[:4], dtracer=z[4])
    a = float(np.sum(d_only[4] * y[4]))
    ad = m.adj_run(*z[:4], ltracer=y[4])
    b = _dot(x[:4], ad[:4])
    assert abs(a) > 0.0 and abs(a - b) <= 1e-10 * abs(a), ("dynamics -> tracer", a, b)


def test_dynamics_unchanged_by_carrying_tracers(window):
    grid, m, _, _ = window
    print(f"case tracer TL/AD leaves the dynamics as without tracers, {m.scheme}", flush=True)
    rng = np.random.default_rng(22)
    x, y = _vectors(grid, rng)
    plain = m.tl_run(*x[:4])
    withtr = m.tl_run(*x[:4], dtracer=x[4])
    for j in range(4):
        assert np.array_equal(plain[j], withtr[j]), j
    # the AD kernels scatter with atomic adds, whose order varies from run
    # to run: equal to rounding
    plain = m.adj_run(*y[:4])
    withtr = m.adj_run(*y[:4], ltracer=np.zeros_like(y[4]))
    for j in range(4):
        assert np.abs(plain[j] - withtr[j]).max() <= 1e-13 * np.abs(plain[j]).max(), j
    assert not np.any(withtr[4])


def test_dot_product_control_with_tracers(window):
    grid, m, c, tr = window
    print(f"case tracer TL/AD dot product, initial state map, {m.scheme}", flush=True)
    rng = np.random.default_rng(23)
    mask = grid.array("grid_center_mask") == 1.0
    s = (grid.iim, grid.jjm, 10)
    dc = [np.where(mask, np.sqrt(np.mean(v[mask] ** 2)) * rng.standard_normal(s), 0.0) for v in c]
    dtr = 0.01 * rng.standard_normal(tr.shape)
    _, y = _vectors(grid, rng)
    for init_only in (True, False):
        tx = m.tl_run(control=dc, init_only=init_only, dtracer=dtr)
        g = m.adj_run(*y[:4], control=True, init_only=init_only, ltracer=y[4])
        a = _dot(tx, y)
        b = _dot(dc, g[:3]) + float(np.sum(dtr * g[3]))
        assert abs(a - b) <= 1e-10 * abs(a), (init_only, a, b)


def test_tracer_gradient_against_finite_differences(window):
    grid, m, c, tr = window
    print(f"case tracer gradient vs central differences, {m.scheme}", flush=True)
    area = grid.array("area")
    mask = grid.array("grid_center_mask") == 1.0
    p = grid.array("p")
    cen = np.array([np.cos(0.2) * np.cos(2.0), np.cos(0.2) * np.sin(2.0), np.sin(0.2)])
    dist = np.arccos(np.clip(np.einsum("k...,k->...", p / np.linalg.norm(p, axis=0), cen), -1, 1)) * S.REF_RADIUS
    cap = mask & (dist < 2000e3)
    assert cap.sum() > 10

    def J(ctl, t0):
        mm = _model(grid, m.scheme)
        mm.init_vorticity(*ctl, tracer=t0)
        mm.step(NSTEPS)
        h, t = mm.field("h"), mm.field("tracer")
        return float(np.sum((area * h * t[:, :, 0, :])[cap])), h, t

    _, hN, tN = J(c, tr)
    assert np.array_equal(m.field("h"), hN) and np.array_equal(m.field("tracer"), tN)
    lt = np.zeros_like(tr)
    lt[:, :, 0, :] = np.where(cap, area * hN, 0.0)
    gz, gd, gh, gt = m.adj_run(np.where(cap, area * tN[:, :, 0, :], 0.0), control=True, ltracer=lt)
    if m.scheme == "fct":
        below, edges = m.tracer_limiter_stats()
        print(f"  frozen tracer limiter coefficients below 1: {below} of {edges} edge-steps", flush=True)
    rng = np.random.default_rng(24)
    s = (grid.iim, grid.jjm, 10)
    eps = 1e-6
    for r in range(6):
        # directions 4 and 5 perturb the initial tracers alone
        d = [np.where(mask, np.sqrt(np.mean(v[mask] ** 2)) * rng.standard_normal(s), 0.0) * (r < 4) for v in c]
        dt0 = np.where(mask[:, :, None, :], 0.1 * rng.standard_normal(tr.shape), 0.0)
        jp = J([a + eps * b for a, b in zip(c, d)], tr + eps * dt0)[0]
        jm = J([a - eps * b for a, b in zip(c, d)], tr - eps * dt0)[0]
        fd = (jp - jm) / (2 * eps)
        ad = float(np.sum(gz * d[0]) + np.sum(gd * d[1]) + np.sum(gh * d[2]) + np.sum(gt * dt0))
        print(f"  direction {r}: fd {fd:.12e} ad {ad:.12e} rel {abs(fd - ad) / abs(fd):.1e}", flush=True)
        assert abs(fd - ad) <= 1e-7 * abs(fd), (r, fd, ad)


def test_tracer_entry_points_refuse():
    print("case tracer TL/AD refusals", flush=True)
    grid = S.Grid(3)
    m = S.Model(grid, ntracer=0, dt=400.0, grav=S.REF_GRAV)
    c, _ = _control(grid)
    m.init_vorticity(*c)
    m.adj_forward(2, 1)
    s = (grid.iim, grid.jjm, 10)
    z = np.zeros(s)
    with pytest.raises(S.SwmError):   # no tracers in this context
        m.tl_run(z, z, z, np.zeros((2, 2) + s), dtracer=np.zeros((grid.iim, grid.jjm, 0, 10)))
    m.close()
