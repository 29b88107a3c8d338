"""Tangent-linear (TL) and adjoint (AD) model of a window of FP64 steps
(swm_tl_run / swm_adj_run), on the device, for each advection scheme:
centred, upstream, and flux-corrected transport with the limiter's
coefficients frozen at the base trajectory.

The linearised model is the derivative of the forward as it runs (each
step's solve is its V-cycles from the previous solution, the default of
swm_set_linearisation); the forward here solves to a relative residual of
1e-13.  The other choice (the solves taken as converged inverses) is
checked by a dot-product test of its own.

Checks, as the linearised model's own references:
  - repeated runs on the same input agree: TL bit for bit, AD to 1e-12 of
    its largest value (its atomic adds change order);
  - dot-product tests <TL x, y> = <x, AD y> over every slot of every state
    field (h, eta, div, u), for the window alone, for the initial state map
    alone (vorticity, divergence, thickness -> state) and for both, to
    1e-10 relative;
  - the gradient of J = sum over a spherical cap of area * (h - 5000)^2
    after the window, with respect to initial vorticity, divergence and
    thickness, against central finite differences of the nonlinear model
    (eps = 1e-6 relative, seeded random directions), to 1e-7 relative.
The window has a partial last segment (8 steps, snapshots every 3), so the
recomputation from snapshots is covered too.  With FCT the finite-difference
check holds as well because the mass limiter (which keeps the mass positive)
is inactive on these fields: every frozen coefficient is 1, and the frozen
linearisation is the derivative.
"""
import numpy as np
import pytest

import swm_mi355x as S

pytestmark = pytest.mark.gpu

G, NSTEPS, EVERY = 4, 8, 3
DT = 200.0 * 2 ** (5 - G)


def _control(grid):
    p = grid.array("p")
    x, y, z = p[0], p[1], p[2]
    zeta = 1e-5 * (z + 0.5 * x * y + 0.3 * y * y)
    delta = 1e-6 * (x * z - 0.2 * y)
    h = 5000.0 + 100.0 * (x + 0.3 * z * z)
    return [np.asfortranarray(a) for a in (zeta, delta, h)]


SCHEMES = ("centered", "fct", "upwind")


def _model(grid, adv="centered"):
    m = S.Model(grid, ntracer=1, dt=DT, grav=S.REF_GRAV, advection=adv)
    m.set_solver(tol=1e-13, relative=True, max_cycles=60)
    return m


@pytest.fixture(scope="module", params=SCHEMES)
def window(request):
    grid = S.Grid(G)
    m = _model(grid, request.param)
    c = _control(grid)
    m.init_vorticity(*c)
    m.adj_forward(NSTEPS, EVERY)
    m.scheme = request.param
    return grid, m, c


def _dot(a, b):
    return sum(float(np.sum(x * y)) for x, y in zip(a, b))


def _state_vectors(grid, rng):
    s = (grid.iim, grid.jjm, 10)
    x = [rng.standard_normal(s), 1e-6 * rng.standard_normal(s), 1e-7 * rng.standard_normal(s),
         0.1 * rng.standard_normal((2, 2) + s)]
    y = [rng.standard_normal(s), 1e6 * rng.standard_normal(s), 1e7 * rng.standard_normal(s),
         10 * rng.standard_normal((2, 2) + s)]
    return x, y


def test_dot_product_window(window):
    grid, m, _ = window
    print(f"case TL/AD dot product, window, {m.scheme}", flush=True)
    rng = np.random.default_rng(11)
    x, y = _state_vectors(grid, rng)
    a = _dot(m.tl_run(*x), y)
    b = _dot(x, m.adj_run(*y))
    assert abs(a - b) <= 1e-10 * abs(a), (a, b)
    # field by field (input k, output j)
    for k in range(4):
        xk = [v if i == k else np.zeros_like(v) for i, v in enumerate(x)]
        tk = m.tl_run(*xk)
        for j in range(4):
            yj = [v if i == j else np.zeros_like(v) for i, v in enumerate(y)]
            aj = float(np.sum(tk[j] * y[j]))
            bj = float(np.sum(x[k] * m.adj_run(*yj)[k]))
            assert abs(aj - bj) <= 1e-10 * max(abs(aj), 1e-300), (k, j, aj, bj)


def test_repeated_runs_agree(window):
    # every run clears the linearised state on the device before it copies
    # its inputs in; the same input gives the same TL output bit for bit, and
    # the same AD output to round-off (its kernels scatter with atomic adds,
    # whose order varies)
    grid, m, _ = window
    print(f"case TL/AD repeated runs, {m.scheme}", flush=True)
    rng = np.random.default_rng(5)
    x, y = _state_vectors(grid, rng)
    for k in range(4):
        xk = [v if i == k else np.zeros_like(v) for i, v in enumerate(x)]
        yk = [v if i == k else np.zeros_like(v) for i, v in enumerate(y)]
        t0, a0 = m.tl_run(*xk), m.adj_run(*yk)
        for _ in range(3):
            t1, a1 = m.tl_run(*xk), m.adj_run(*yk)
            for j in range(4):
                assert np.array_equal(t0[j], t1[j]), (k, j, "tl")
                scale = max(float(np.max(np.abs(a0[j]))), 1e-300)
                assert float(np.max(np.abs(a1[j] - a0[j]))) <= 1e-12 * scale, (k, j, "adj")


def test_dot_product_initial_state_map(window):
    grid, m, c = window
    print(f"case TL/AD dot product, initial state map, {m.scheme}", flush=True)
    rng = np.random.default_rng(12)
    mask = grid.array("grid_center_mask") == 1.0
    s = (grid.iim, grid.jjm, 10)
    dc = [np.where(mask, np.sqrt(np.mean(v[mask] ** 2)) * rng.standard_normal(s), 0.0) for v in c]
    _, y = _state_vectors(grid, rng)
    a = _dot(m.tl_run(control=dc, init_only=True), y)
    b = _dot(dc, m.adj_run(*y, init_only=True))
    assert abs(a - b) <= 1e-10 * abs(a), (a, b)
    a = _dot(m.tl_run(control=dc), y)
    b = _dot(dc, m.adj_run(*y, control=True))
    assert abs(a - b) <= 1e-10 * abs(a), (a, b)


def test_gradient_against_finite_differences(window):
    grid, m, c = window
    print(f"case adjoint gradient vs central differences, {m.scheme}", flush=True)
    area = grid.array("area")
    mask = grid.array("grid_center_mask") == 1.0
    p = grid.array("p")
    cen = np.array([np.cos(0.6) * np.cos(1.0), np.cos(0.6) * np.sin(1.0), np.sin(0.6)])
    dist = np.arccos(np.clip(np.einsum("k...,k->...", p / np.linalg.norm(p, axis=0), cen), -1, 1)) * S.REF_RADIUS
    cap = mask & (dist < 1500e3)
    assert cap.sum() > 10

    def J(ctl):
        mm = _model(grid, m.scheme)
        mm.init_vorticity(*ctl)
        mm.step(NSTEPS)
        h = mm.field("h")
        return float(np.sum((area * (h - 5000.0) ** 2)[cap])), h

    _, hN = J(c)
    # the window's final thickness is the nonlinear model's
    assert np.array_equal(m.field("h"), hN)
    gz, gd, gh = m.adj_run(np.where(cap, 2.0 * area * (hN - 5000.0), 0.0), control=True)
    rng = np.random.default_rng(13)
    s = (grid.iim, grid.jjm, 10)
    eps = 1e-6
    for r in range(4):
        d = [np.where(mask, np.sqrt(np.mean(v[mask] ** 2)) * rng.standard_normal(s), 0.0) for v in c]
        jp, _ = J([a + eps * b for a, b in zip(c, d)])
        jm, _ = J([a - eps * b for a, b in zip(c, d)])
        fd = (jp - jm) / (2 * eps)
        ad = float(np.sum(gz * d[0]) + np.sum(gd * d[1]) + np.sum(gh * d[2]))
        assert abs(fd - ad) <= 1e-7 * abs(fd), (r, fd, ad)


def test_window_errors():
    grid = S.Grid(3)
    c = _control(grid)
    m = S.Model(grid, ntracer=1, dt=800.0, grav=S.REF_GRAV, precision=S.SWM_MIXED)   # FP64 only
    m.init_synthetic(lmax=8)
    with pytest.raises(S.SwmError):
        m.adj_forward(4, 2)
    m = S.Model(grid, ntracer=1, dt=800.0, grav=S.REF_GRAV, advection="centered")
    with pytest.raises(S.SwmError):
        m.tl_run(control=c)   # no window yet
    m.init_vorticity(*c)
    m.step(1)
    with pytest.raises(S.SwmError):
        m.adj_forward(4, 2)   # not at step 0


def test_torch_function_gives_the_adjoint_gradient(window):
    # the gradient harness: J built in torch on WindowForecast's output
    import torch
    from swm_mi355x import adjoint as A
    grid, w, c = window
    area = grid.array("area")
    cap = A.cap_mask(grid, 0.6, 1.0, 1500e3, S.REF_RADIUS)
    m = _model(grid, w.scheme)
    z, d, h = (torch.from_numpy(np.ascontiguousarray(v)).requires_grad_(True) for v in c)
    J = A.cap_objective(A.WindowForecast.apply(z, d, h, m, NSTEPS, EVERY), area, cap, 5000.0)
    J.backward()
    hN = m.field("h")
    assert abs(J.item() - A.cap_objective(hN, area, cap, 5000.0)) <= 1e-12 * abs(J.item())
    gz, gd, gh = m.adj_run(np.where(cap, 2.0 * area * (hN - 5000.0), 0.0), control=True)
    # the same adjoint; atomic accumulation orders differ between runs
    for t, g in ((z, gz), (d, gd), (h, gh)):
        assert np.linalg.norm(t.grad.numpy() - g) <= 1e-12 * np.linalg.norm(g)
    mask = grid.array("grid_center_mask") == 1.0
    res = A.finite_difference_check(lambda: _model(grid, w.scheme), c, (gz, gd, gh),
                                    lambda hh: A.cap_objective(hh, area, cap, 5000.0), NSTEPS, 2, 14, mask)
    assert all(r[2] <= 1e-7 for r in res), res


def test_dot_product_with_active_limiter():
    # a thin layer (h down to ~1 m beside 5000 m) makes the positivity limiter
    # of the mass FCT act at the layer's edge, so frozen coefficients below 1
    # enter the linearisation; the adjoint is still that of the TL to 1e-10
    print("case TL/AD dot product, FCT with an active limiter", flush=True)
    grid = S.Grid(G)
    p = grid.array("p")
    zeta, delta, _ = _control(grid)
    h = 1.0 + 4999.0 * 0.5 * (1.0 + np.tanh((p[0] - 0.3) / 0.08))
    m = _model(grid, "fct")
    m.init_vorticity(zeta, delta, np.asfortranarray(h))
    m.adj_forward(NSTEPS, EVERY)
    assert np.isfinite(m.field("h")).all() and m.field("h")[grid.array("grid_center_mask") == 1.0].min() > 0.0
    rng = np.random.default_rng(21)
    x, y = _state_vectors(grid, rng)
    a = _dot(m.tl_run(*x), y)
    b = _dot(x, m.adj_run(*y))
    assert abs(a - b) <= 1e-10 * abs(a), (a, b)


@pytest.mark.parametrize("scheme", ["fct", "centered"])
def test_converged_solve_linearisation_dot_product(scheme):
    print(f"case TL/AD dot product, converged-solve linearisation, {scheme}", flush=True)
    grid = S.Grid(G)
    m = _model(grid, scheme)
    m.set_linearisation(False)
    m.init_vorticity(*_control(grid))
    m.adj_forward(NSTEPS, EVERY)
    rng = np.random.default_rng(31)
    x, y = _state_vectors(grid, rng)
    a = _dot(m.tl_run(*x), y)
    b = _dot(x, m.adj_run(*y))
    assert abs(a - b) <= 1e-10 * abs(a), (a, b)


def test_production_forward_gradient_fourth_order():
    # the reference criterion (2-3 V-cycles per solve) and the exact
    # linearisation; with 3 fixed cycles per solve the forward is smooth, and
    # fourth-order differences at eps = 1e-4 check the gradient to 1e-8
    print("case TL/AD production forward, fixed cycles, fourth-order differences", flush=True)
    from swm_mi355x import adjoint as A
    grid = S.Grid(G)
    area = grid.array("area")
    mask = grid.array("grid_center_mask") == 1.0
    cap = A.cap_mask(grid, 0.6, 1.0, 1500e3, S.REF_RADIUS)
    c = _control(grid)

    def model():
        mm = S.Model(grid, ntracer=1, dt=DT, grav=S.REF_GRAV, advection="fct")
        mm.set_fixed_cycles(3, 0)
        return mm

    m = model()
    m.init_vorticity(*c)
    m.adj_forward(NSTEPS, EVERY)
    hN = m.field("h")
    grad = m.adj_run(np.where(cap, 2.0 * area * (hN - 5000.0), 0.0), control=True)
    res = A.finite_difference_check(model, c, grad, lambda hh: A.cap_objective(hh, area, cap, 5000.0), NSTEPS, 4, 17,
                                    mask, eps=1e-4, order=4)
    assert all(r[2] <= 1e-8 for r in res), res
