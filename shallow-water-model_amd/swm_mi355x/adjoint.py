"""Gradient harness of the linearised model (swm_tl_run / swm_adj_run).

``WindowForecast`` is a torch.autograd.Function whose forward runs the FP64
model (centred advection) over a window from the initial state given by
relative vorticity, divergence and thickness, and returns the thickness
after the window; its backward is the hand-written adjoint of the window
and of the initial state map.  Torch only carries the tensors here; the
computation is the library's.

``cap_objective`` is the scalar objective of the gradient checks: the
area-weighted sum of (h - h_ref)^2 over a spherical cap.
"""
import numpy as np
import torch

from . import Model


class WindowForecast(torch.autograd.Function):
    """h (iim,jjm,10) after `nsteps` steps from the state of
    Model.init_vorticity(zeta, delta, h0); `model` is an FP64 centred-advection
    Model, left at the window's end."""

    @staticmethod
    def forward(ctx, zeta, delta, h0, model, nsteps, every):
        model.init_vorticity(zeta.detach().cpu().numpy(), delta.detach().cpu().numpy(),
                             h0.detach().cpu().numpy())
        model.adj_forward(nsteps, every)
        ctx.model = model
        return torch.from_numpy(np.ascontiguousarray(model.field("h")))

    @staticmethod
    def backward(ctx, gh):
        gz, gd, gh0 = ctx.model.adj_run(gh.detach().cpu().numpy(), control=True)
        return (torch.from_numpy(np.ascontiguousarray(gz)), torch.from_numpy(np.ascontiguousarray(gd)),
                torch.from_numpy(np.ascontiguousarray(gh0)), None, None, None)


def cap_mask(grid, lat, lon, radius_m, sphere_radius):
    """Real cells within radius_m (great-circle distance) of (lat, lon)."""
    p = grid.array("p")
    cen = np.array([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)])
    d = np.arccos(np.clip(np.einsum("k...,k->...", p / np.linalg.norm(p, axis=0), cen), -1.0, 1.0))
    return (grid.array("grid_center_mask") == 1.0) & (d * sphere_radius < radius_m)


def cap_objective(h, area, cap, h_ref):
    """J = sum over the cap of area * (h - h_ref)^2 (torch or numpy)."""
    if isinstance(h, torch.Tensor):
        a = torch.from_numpy(np.where(cap, area, 0.0))
        return torch.sum(a * (h - h_ref) ** 2)
    return float(np.sum((area * (h - h_ref) ** 2)[cap]))


def forecast(model_factory, control, nsteps):
    """The nonlinear model's h after nsteps from the control (zeta, delta,
    h0): the function whose derivative the adjoint gives."""
    m = model_factory()
    m.init_vorticity(*control)
    m.step(nsteps)
    return m.field("h")


def finite_difference_check(model_factory, control, gradient, objective, nsteps, ndir, seed, mask, eps=1e-6,
                            order=2):
    """Central differences of objective(h after nsteps) along ndir seeded
    random directions in control space (each component scaled to its RMS
    over the real cells, mask, eps relative), against <gradient, direction>.
    order 2: (J(e) - J(-e)) / 2e; order 4: (8 (J(e) - J(-e)) - (J(2e) -
    J(-2e))) / 12e, whose truncation error is O(e^4), so that a larger eps
    (less FP64 rounding noise of J divided by e) can be used.
    Returns a list of (fd, adjoint, relative difference)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(ndir):
        d = []
        for c in control:
            rms = np.sqrt(np.mean(c[mask] ** 2))
            d.append(np.where(mask, rms * rng.standard_normal(c.shape), 0.0))

        def J(t):
            return objective(forecast(model_factory, [a + t * b for a, b in zip(control, d)], nsteps))

        if order == 4:
            fd = (8.0 * (J(eps) - J(-eps)) - (J(2.0 * eps) - J(-2.0 * eps))) / (12.0 * eps)
        else:
            fd = (J(eps) - J(-eps)) / (2.0 * eps)
        ad = float(sum(np.sum(g * b) for g, b in zip(gradient, d)))
        out.append((fd, ad, abs(fd - ad) / max(abs(fd), 1e-300)))
    return out


__all__ = ["WindowForecast", "cap_mask", "cap_objective", "forecast", "finite_difference_check", "Model"]
