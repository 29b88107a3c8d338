"""Benchmark of the MI355X shallow-water dynamical core.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--g 9] [--error-steps N]

One "step" is one full model time step over the whole sphere:

- FCT transport of mass and one tracer;
- the ABM3 vorticity/divergence update;
- the multigrid Poisson solve of streamfunction and velocity potential to the reference's criterion;
- velocity, kinetic energy and halo fills.

The workload is the level-9 icosahedral grid (2,621,442 cells) with seeded
synthetic fields of harmonic degree 1..160, built on the device. Inputs are
resident in HBM before the timed region starts.

**Time step.** dt = 200·2^(5−g) s, i.e. 12.5 s at g=9. That is the
reference's own choice (200 s at 10242 cells, `swm/run.0010242/data/namelist`),
scaled with the grid spacing. The configured dt = 450·2^(5−g) s violates the
scheme's stability limit on these fields: at g=9 the state blows up and the
elliptic solve fails its criterion at step 25, as the reference does at g=3.
The V-cycles per step depend on dt (the warm start is older by one step), so
the line reports them, and a scan over dt = 25, 20, 16 s reports the largest
step at which the g=9 day completes, with its cell-steps/s.

**Forms timed on the same initial state.** K steps after W warm-up steps:

- the mixed-precision step (FP32 state and stencils, FP64 defect correction): the headline `value`;
- the FP64 step (the reference's arithmetic);
- the mixed step on the FP64 step's solver schedule (the last solution as the warm start, two cycles before the first test), which separates the gain of the precision from that of the schedule.

The mixed and FP64 forms then continue to one simulated day (6912 steps at
g=9); the whole day is timed and its V-cycles per step reported. The mixed
path's error against the FP64 path is reported per field (area-weighted
relative L2 and relative max), also with both forms solved to mean(r^2) <=
1e-24, where only the precision differs.

**Level 8** (655362 cells, dt = 25 s): the FP64 step over one simulated day
(3456 steps), and a microbench of single operators on its resident state:
div_of_an_edge_flux, curl of a corner vector, the Laplacian and the face
halo fill, in microseconds and GB/s against the 8 TB/s HBM peak.

**The JSON line carries:**

- `roofline`: the hottest kernel of the headline path (the finest-level multigrid smoother), timed with HIP events on the model's stream during the timed steps.
- `step_roofline`: the whole step against the canonical bytes per cell-step.
- `cpu_baseline`: the oracle, the C restatement of the reference FP64 step, on the host's cores (OpenMP over the panels) and on one core, over a bounded sample of steps of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "shallow-water-model_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "cell-steps/s at g=9, mixed vs FP64 at stated err; HBM roofline %; adj/fwd ratio"
RADIUS, GRAV, OMEGA = 6.371e6, 9.81, 7.292e-5
LMAX = {5: 20, 6: 40, 7: 40, 8: 80, 9: 160}


def step_bytes(precision, ntracer, vcycles_per_step, checks_per_step):
    """Canonical algorithmic bytes per cell-step: compulsory traffic of the
    state (S words: h 2, eta 2, div 2, u 8, ke 2, psi/chi 4, ABM3 histories 6,
    FCT edge-flux history 9, 4 per tracer) and the metrics (M = 67 words)
    once, plus per V-cycle 80 words (four finest sweeps of 13 words,
    restriction 3.25, prolongation 4.5, times 4/3 over the levels) and per
    residual check 11 words.  Mixed: FP32 state and metrics, FP64
    psi/chi masters (16 B), the first FP64 defect from the extrapolated
    start (120 B: two solutions read, the start written, FP64 weights, FP32
    rhs and defect), then per correction cycle an FP32 V-cycle (320 B) and
    the FP64 correction + defect (128 B), which also yields the sums of the
    convergence test; mean removal and mass sum 24 B."""
    S = 35 + 4 * ntracer
    M = 67
    if precision == "fp64":
        return 8 * (S + M + 5) + 8 * (vcycles_per_step * 80 + checks_per_step * 11)
    n_o = vcycles_per_step
    return 4 * (S + M) + 16 + 120 + n_o * (128 + 320) + 24


def rel_errors(grid, a, b):
    import numpy as np
    mask = grid.array("grid_center_mask") == 1.0
    area = grid.array("area")
    out = {}
    for k in ("h", "eta", "div", "tracer"):
        x, y = a[k], b[k]
        mm, w = mask, area
        if x.ndim == 4:
            mm = np.broadcast_to(mask[:, :, None, :], x.shape)
            w = np.broadcast_to(area[:, :, None, :], x.shape)
        d, ww, xv = (x - y)[mm], w[mm], x[mm]
        out[k] = {"rel_l2": float(np.sqrt(np.sum(ww * d * d) / np.sum(ww * xv * xv))),
                  "rel_linf": float(np.abs(d).max() / np.abs(xv).max())}
    return out


def timed(model, warmup, steps):
    model.synchronize()
    tw = time.perf_counter()
    model.step(warmup)
    model.synchronize()
    warm = time.perf_counter() - tw
    s0, v0, c0 = model.solver_stats()
    # every 7th pair of back-to-back finest-level smoother launches is timed
    # with HIP events (the event records would cost the step ~3 us each if
    # every pair had them; 7 is prime to the pairs of a step, 1 mixed and 6
    # FP64, so the sample covers every position in the V-cycles)
    model.profile(True, every=7)
    t1 = time.perf_counter()
    model.step(steps)
    model.synchronize()
    wall = time.perf_counter() - t1
    model.profile(False)
    s1, v1, c1 = model.solver_stats()
    dev_ms, k_ms, k_bytes, k_n = model.timing()
    return {"wall": wall, "warm_wall": warm, "device_ms": dev_ms, "kernel_ms": k_ms, "kernel_bytes": k_bytes, "kernel_launches": k_n,
            "vcycles_per_step": (v1 - v0) / max(1, s1 - s0), "checks_per_step": (c1 - c0) / max(1, s1 - s0)}


def adjoint_bench(S, g=7, hours=6.0, every=16, ndir=8, advection="fct"):
    """The adjoint form: FP64, the reference's configured advection (FCT,
    linearised with the limiter's coefficients frozen at the base
    trajectory), a window of `hours` at level g (dt = 200 * 2^(5-g) s),
    snapshots every `every` steps.  The forward is the production one: the
    reference's elliptic criterion (2-3 V-cycles per solve); the linearised
    model is its exact derivative (each solve its V-cycles from the previous
    solution).  Control: relative vorticity, divergence and thickness of the
    seeded synthetic state (degrees 1..40); objective: the area-weighted sum
    of (h - 5000)^2 over a cap of radius 1500 km.  Times the forward, the
    adjoint (segment recomputation + adjoint steps + initial-state map) and
    the tangent-linear run; checks the adjoint by a dot-product test and the
    gradient by fourth-order central differences along ndir seeded
    directions of the same forward with 3 V-cycles per solve (a smooth
    function of the control; eps = 1e-4)."""
    import numpy as np
    from swm_mi355x import adjoint as A
    dt = 200.0 * 2.0 ** (5 - g)
    nsteps = int(round(hours * 3600.0 / dt))
    grid = S.Grid(g, radius=RADIUS, omega=OMEGA)

    def model(fixed=0):
        m = S.Model(grid, ntracer=1, dt=dt, grav=GRAV, advection=advection)
        if fixed:
            m.set_fixed_cycles(fixed, 0)
        return m

    syn = model()
    syn.init_synthetic(seed_index=3, lmax=40)
    control = [np.asfortranarray(syn.field(n)) for n in ("relative", "div", "h")]
    mask = grid.array("grid_center_mask") == 1.0
    area = grid.array("area")
    cap = A.cap_mask(grid, np.pi / 6, -np.pi / 2, 1500e3, RADIUS)
    s = (grid.iim, grid.jjm, 10)

    ref = model()
    ref.init_vorticity(*control)
    ref.synchronize()
    t0 = time.perf_counter()
    ref.step(nsteps)
    ref.synchronize()
    t_fwd = (time.perf_counter() - t0) * 1e3
    m = model()
    m.init_vorticity(*control)
    m.synchronize()
    t0 = time.perf_counter()
    m.adj_forward(nsteps, every)
    m.synchronize()
    t_win = (time.perf_counter() - t0) * 1e3
    hN = m.field("h")
    lh = np.where(cap, 2.0 * area * (hN - 5000.0), 0.0)
    m.adj_run(lh, control=True)   # first call: the coarse cycle's transpose is formed
    t0 = time.perf_counter()
    m.adj_run(lh, control=True)
    t_adj = (time.perf_counter() - t0) * 1e3
    rng = np.random.default_rng(7)
    dc = [np.where(mask, np.sqrt(np.mean(c[mask] ** 2)) * rng.standard_normal(s), 0.0) for c in control]
    t0 = time.perf_counter()
    tl = m.tl_run(control=dc)
    t_tl = (time.perf_counter() - t0) * 1e3
    y = [rng.standard_normal(s), 1e6 * rng.standard_normal(s), 1e7 * rng.standard_normal(s),
         10 * rng.standard_normal((2, 2) + s)]
    a = sum(float(np.sum(u * v)) for u, v in zip(tl, y))
    b = sum(float(np.sum(u * v)) for u, v in zip(dc, m.adj_run(*y, control=True)))
    below, edges = m.limiter_stats()
    # the same window with the tracer carried (tracer FCT, frozen limiter)
    ts = (grid.iim, grid.jjm, 1, 10)
    t0 = time.perf_counter()
    m.adj_run(lh, control=True, ltracer=np.zeros(ts))
    t_adj_tr = (time.perf_counter() - t0) * 1e3
    xt = 0.01 * rng.standard_normal(ts)
    yt = 100.0 * rng.standard_normal(ts)
    t0 = time.perf_counter()
    tlt = m.tl_run(control=dc, dtracer=xt)
    t_tl_tr = (time.perf_counter() - t0) * 1e3
    gtr = m.adj_run(*y, control=True, ltracer=yt)
    a_tr = sum(float(np.sum(u * v)) for u, v in zip(tlt, y + [yt]))
    b_tr = sum(float(np.sum(u * v)) for u, v in zip(dc + [xt], gtr))
    # the gradient check on the smooth forward (3 V-cycles per solve)
    mf = model(3)
    mf.init_vorticity(*control)
    mf.adj_forward(nsteps, every)
    gf = mf.adj_run(np.where(cap, 2.0 * area * (mf.field("h") - 5000.0), 0.0), control=True)
    fd = A.finite_difference_check(lambda: model(3), control, gf, lambda h: A.cap_objective(h, area, cap, 5000.0),
                                   nsteps, ndir, 8, mask, eps=1e-4, order=4)
    scheme = {"fct": "FCT (frozen limiter)", "centered": "centred", "upwind": "upstream"}[advection]
    return {"workload": f"icosahedral level {g} ({grid.max_ig} cells), FP64 {scheme} advection, dt={dt} s, "
                        f"{hours:g} h window ({nsteps} steps), snapshots every {every} steps with recomputation; "
                        f"J = sum over a 1500 km cap of area (h - 5000)^2, gradient with respect to initial "
                        f"vorticity, divergence and thickness",
            "g": g, "steps": nsteps, "checkpoint_every": every,
            "forward_ms": t_fwd, "window_forward_ms": t_win, "adjoint_ms": t_adj, "tl_ms": t_tl,
            "adj_over_fwd": t_adj / t_fwd, "tl_over_fwd": t_tl / t_fwd,
            "vcycles_per_solve": ref.solver_stats()[1] / max(1, ref.solver_stats()[0]),
            "dot_product_rel": abs(a - b) / abs(a),
            "limiter_active_edge_steps": below, "edge_steps": edges,
            "with_tracer": {"adjoint_ms": t_adj_tr, "tl_ms": t_tl_tr, "adj_over_fwd": t_adj_tr / t_fwd,
                            "dot_product_rel": abs(a_tr - b_tr) / abs(a_tr),
                            "note": "the window with its tracer carried through the linearised model "
                                    "(swm_tl_run_tracer / swm_adj_run_tracer; tracer FCT with its limiter frozen)"},
            "fd_rel_errors": [r[2] for r in fd],
            "note": "forward at the reference criterion (mean(r^2) <= 1e-20); the adjoint is the exact transpose of "
                    "that forward's own cycles (recomputation included in adjoint_ms); finite differences: "
                    "fourth order, eps = 1e-4, on the forward with 3 V-cycles per solve"}


def ensemble_bench(S, hours=12.0, n6=96, n7=32, tune_rounds=2, eager_steps=12):
    """The ensemble form: n6 members at level 6 (dt = 100 s) and n7 members at
    level 7 (dt = 50 s, two substeps per level-6 step), mixed precision, one
    context per level with every member in the same launches, both advanced
    together in captured HIP graphs with the autotuner's launch choice, for
    `hours` simulated hours.  Member m: the synthetic construction (harmonic
    degrees 1..40, seed index 4) plus a 1 % vorticity perturbation from
    stream m.  The mixed solve runs a fixed number of correction cycles: the
    smallest of 1, 2, 3 with which every solve of a 24-step calibration run
    meets the reference criterion; the solves of the timed run are recorded
    on the device and reported.  Error: each member against the same member
    run in FP64 (FP64 contexts with the reference's convergence test), and
    the ensemble mean against the FP64 ensemble mean."""
    import numpy as np
    grids = {6: S.Grid(6, radius=RADIUS, omega=OMEGA), 7: S.Grid(7, radius=RADIUS, omega=OMEGA)}
    spec = [(6, n6, 1, 0), (7, n7, 2, n6)]   # level, members, substeps, first member seed

    def group(g, n, seed0, precision, fixed=0):
        m = S.Model(grids[g], ntracer=1, dt=200.0 * 2.0 ** (5 - g), grav=GRAV, precision=precision, members=n)
        if fixed:
            m.set_fixed_cycles(fixed, 8)
        for k in range(n):
            m.select(k)
            m.init_synthetic_member(seed0 + k, 0.01, seed_index=4, lmax=40)
        return m

    def tracks(groups):
        out = []
        for m in groups:
            for k in range(m.members):
                m.select(k)
                out.append(m.solver_track())
        return out

    t0 = time.perf_counter()
    fixed, groups, ens = None, None, None
    for ncyc in (1, 2, 3):
        groups = [group(g, n, s0, S.SWM_MIXED, ncyc) for g, n, _, s0 in spec]
        ens = S.Ensemble(groups, [sub for _, _, sub, _ in spec])
        ens.step(24)
        if all(t[2] == 0 for t in tracks(groups)) or ncyc == 3:
            fixed = ncyc
            break
        ens.close()
        for m in groups:
            m.close()
    # the timed run starts from the initial state
    for (g, n, _, s0), m in zip(spec, groups):
        for k in range(n):
            m.select(k)
            m.init_synthetic_member(s0 + k, 0.01, seed_index=4, lmax=40)
    t_setup = time.perf_counter() - t0
    trials = ens.tune(rounds=tune_rounds)
    choice = ens.choice()
    nsteps = int(round(hours * 3600.0 / 100.0))
    # the first 12 ensemble steps capture the graphs (six start-up phases,
    # then the six periodic ones); the rest replay them and are timed
    nwarm = 12
    ens.step(nwarm)
    t1 = time.perf_counter()
    ens.step(nsteps - nwarm)
    wall = time.perf_counter() - t1
    dev_ms = ens.choice()["ms_per_step"]
    tr = tracks(groups)
    member_steps = sum(n * sub for _, n, sub, _ in spec) * (nsteps - nwarm)
    cell_steps = sum(n * sub * grids[g].max_ig for g, n, sub, _ in spec) * (nsteps - nwarm)
    # error against FP64 runs of the same members
    ref = [group(g, n, s0, S.SWM_FP64) for g, n, _, s0 in spec]
    for (g, n, sub, _), m in zip(spec, ref):
        m.step(sub * nsteps)
    errs = {}
    for (g, n, sub, _), mm, mr in zip(spec, groups, ref):
        grid = grids[g]
        mask = grid.array("grid_center_mask") == 1.0
        area = grid.array("area")[mask]
        per = {k: [] for k in ("h", "eta", "div", "tracer")}
        mean_m = {k: 0.0 for k in per}
        mean_r = {k: 0.0 for k in per}
        for k in range(n):
            mm.select(k)
            mr.select(k)
            for f in per:
                x = mr.field(f)
                y = mm.field(f)
                if x.ndim == 4:
                    x, y = x[:, :, 0, :], y[:, :, 0, :]
                x, y = x[mask], y[mask]
                per[f].append(float(np.sqrt(np.sum(area * (y - x) ** 2) / np.sum(area * x * x))))
                mean_m[f] = mean_m[f] + y / n
                mean_r[f] = mean_r[f] + x / n
        errs[f"g{g}"] = {f: {"member_rel_l2_mean": float(np.mean(v)), "member_rel_l2_max": float(np.max(v)),
                             "ensemble_mean_rel_l2": float(np.sqrt(np.sum(area * (mean_m[f] - mean_r[f]) ** 2)
                                                                  / np.sum(area * mean_r[f] ** 2)))}
                         for f, v in per.items()}
    for m in ref:
        m.close()
    # the same steps one launch at a time (no graph), for the graphs' gain
    for m in groups:
        m.synchronize()
    t2 = time.perf_counter()
    for _ in range(eager_steps):
        for (g, n, sub, _), m in zip(spec, groups):
            m.step(sub)
    for m in groups:
        m.synchronize()
    eager_ms = (time.perf_counter() - t2) * 1e3 / eager_steps
    ens.close()
    for m in groups:
        m.close()
    return {"members": {"g6": n6, "g7": n7}, "dt": {"g6": 100.0, "g7": 50.0}, "substeps": {"g6": 1, "g7": 2},
            "simulated_hours": nsteps * 100.0 / 3600.0, "ensemble_steps": nsteps,
            "member_steps_per_s": member_steps / wall, "cell_steps_per_s": cell_steps / wall,
            "timed_ensemble_steps": nsteps - nwarm,
            "ms_per_ensemble_step": wall * 1e3 / (nsteps - nwarm), "device_ms_per_ensemble_step": dev_ms,
            "eager_ms_per_ensemble_step": eager_ms,
            "graph_speedup_vs_eager": eager_ms / (wall * 1e3 / (nsteps - nwarm)),
            "fixed_cycles": fixed, "fixed_cycles_start": 8,
            "solves_above_criterion": int(sum(t[2] for t in tr)), "max_mean_resid2": max(max(t[0], t[1]) for t in tr),
            "autotuner_choice": {k: choice[k] for k in ("concurrent", "steps_per_graph", "blk_step", "blk_mg")},
            "autotuner_trials": trials, "setup_s": t_setup,
            "error_vs_fp64": errs,
            "note": "relative L2 (area-weighted, real cells) of each member against the same member in FP64 "
                    "(mean and max over members) and of the ensemble mean; dt = 200*2^(5-g) s as the "
                    "single-model form (the configured 450*2^(5-g) s is past the scheme's stability limit)"}


def level8_bench(S, reps=200):
    """Level 8 (655362 cells), FP64, dt = 25 s (the reference's own time step
    at this size, swm/run.0655362/data/namelist): one simulated day (3456
    steps) timed as a whole, then the single-operator microbench on the
    resident state."""
    g, dt = 8, 25.0
    grid = S.Grid(g, radius=RADIUS, omega=OMEGA)
    m = S.Model(grid, ntracer=1, dt=dt, grav=GRAV)
    m.init_synthetic(seed_index=1, lmax=LMAX[g])
    nsteps = int(round(86400.0 / dt))
    m.synchronize()
    t0 = time.perf_counter()
    m.step(nsteps)
    m.synchronize()
    wall = time.perf_counter() - t0
    solves, vcyc, checks = m.solver_stats()
    ops = {}
    for op in S.Model.OPERATOR_BENCH[:4]:
        us, nbytes = m.operator_bench(op, reps)
        gbs = nbytes / (us * 1e-6) / 1e9
        ops[op] = {"us": us, "bytes_per_launch": nbytes, "gbs": gbs, "frac_of_hbm_peak": gbs / HBM_PEAK_GBS}
    out = {"workload": f"icosahedral level {g} ({grid.max_ig} cells), FP64 step, dt={dt} s, FCT mass + 1 tracer, "
                       f"multigrid to mean(resid^2)<=1e-20, one simulated day ({nsteps} steps)",
           "cells": grid.max_ig, "dt": dt, "steps": nsteps, "wall_s": wall,
           "cell_steps_per_s": grid.max_ig * nsteps / wall, "ms_per_step": wall * 1e3 / nsteps,
           "vcycles_per_step": vcyc / max(1, solves),
           "step_roofline_frac": grid.max_ig * nsteps / wall * step_bytes("fp64", 1, vcyc / max(1, solves),
                                                                          checks / max(1, solves)) / 1e9
           / HBM_PEAK_GBS,
           "operators": ops,
           "operators_note": "reps back-to-back launches of the kernel each operator entry point runs, between HIP "
                             "events; algorithmic FP64 bytes per real cell: div_of_an_edge_flux 40 B, curl of a "
                             "corner vector 144 B, Laplacian 72 B; face halo fill 16 B per ghost slot"}
    m.close()
    grid.close()
    return out


def remap_bench(S, grid, dt, lmax, ntracer, warmup, steps, reps=100):
    """Incremental remapping (l_remap, FP64) on the headline grid and time
    step: steps timed after warmup, the FP64 FCT step beside it for the
    ratio, then the remapping's transport and field kernels and one
    application of the dissipation operator on the resident state."""
    import numpy as np
    m = S.Model(grid, ntracer=ntracer, dt=dt, grav=GRAV, advection="remap")
    m.init_synthetic(seed_index=1, lmax=lmax)
    t = timed(m, warmup, steps)
    cells = grid.max_ig
    ops = {}
    m.init_diss_op()
    for op in ("remap_transport", "remap_fields", "dissipation"):
        us, nbytes = m.operator_bench(op, reps)
        gbs = nbytes / (us * 1e-6) / 1e9
        ops[op] = {"us": us, "bytes_per_launch": nbytes, "gbs": gbs, "frac_of_hbm_peak": gbs / HBM_PEAK_GBS}
    h = m.field("h")
    mask = grid.array("grid_center_mask") == 1.0
    finite = bool(np.isfinite(h[mask]).all())
    m.close()
    return {"workload": f"icosahedral level {grid.g} ({cells} cells), FP64 step with incremental remapping of "
                        f"mass and {ntracer} tracer(s) (get_remap_flux), dt={dt} s",
            "steps": steps, "warmup": warmup, "ms_per_step": t["wall"] * 1e3 / steps,
            "cell_steps_per_s": cells * steps / t["wall"], "finite": finite, "kernels": ops,
            "kernels_note": "reps back-to-back launches between HIP events; algorithmic FP64 bytes per real cell: "
                            "transport 1128 + 48 ntracer, fields 288 + 32 ntracer, dissipation 368"}


def dt_scan(S, grid, lmax, ntracer, dts=(25.0, 20.0, 16.0)):
    """The mixed step over one simulated day at each dt, largest first, until
    one completes (a blow-up ends in an elliptic solve that misses its
    criterion)."""
    out = []
    for dt in dts:
        m = S.Model(grid, ntracer=ntracer, dt=dt, grav=GRAV, precision=S.SWM_MIXED)
        m.init_synthetic(seed_index=1, lmax=lmax)
        n = int(round(86400.0 / dt))
        m.synchronize()
        t0 = time.perf_counter()
        done, ok = 0, True
        try:
            while done < n:
                k = min(256, n - done)
                m.step(k)
                done += k
            m.synchronize()
        except S.SwmError:
            ok = False
            done = m.step_count
        wall = time.perf_counter() - t0
        solves, vcyc, _ = m.solver_stats()
        rec = {"dt": dt, "steps": n, "completed": ok, "steps_done": done}
        if ok:
            rec.update({"cell_steps_per_s": grid.max_ig * n / wall, "ms_per_step": wall * 1e3 / n,
                        "vcycles_per_step": vcyc / max(1, solves)})
        out.append(rec)
        m.close()
        if ok:
            break
    return out


def cpu_baseline_run(grid, ntracer, dt, state, steps, threads):
    from oracle.oracle import Oracle, grid_arrays_from, set_threads
    set_threads(threads)
    orc = Oracle(grid_arrays_from(grid), ntracer, dt, GRAV)
    orc.init_state(*state)
    tc = time.perf_counter()
    orc.step(steps)
    t = time.perf_counter() - tc
    orc.close()
    return grid.max_ig * steps / t, t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--g", type=int, default=9)
    ap.add_argument("--ntracer", type=int, default=1)
    ap.add_argument("--error-steps", type=int, default=None,
                    help="steps of the day window and the mixed-vs-FP64 error (default: one simulated day; 0 skips)")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint form (level 7, 6 h window)")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the ensemble form (96 members at level 6 and 32 at level 7, 12 h)")
    ap.add_argument("--no-level8", action="store_true", help="skip the level-8 FP64 day and operator microbench")
    ap.add_argument("--no-dt-scan", action="store_true", help="skip the scan over larger time steps")
    ap.add_argument("--no-remap", action="store_true", help="skip the incremental-remapping step")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the panel-sharded multi-GPU design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import swm_mi355x as S

    g = args.g
    dt = 200.0 * 2.0 ** (5 - g)
    lmax = LMAX.get(g, 20)
    t0 = time.time()
    grid = S.Grid(g, radius=RADIUS, omega=OMEGA)
    cells = grid.max_ig
    print(f"[bench] grid g={g} ({cells} cells) built in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    def model(precision, tol=None, reference_schedule=False):
        m = S.Model(grid, ntracer=args.ntracer, dt=dt, grav=GRAV, precision=precision)
        if tol:
            m.set_solver(tol=tol)
        if reference_schedule:
            m.set_mixed_start(0, 2)
        m.init_synthetic(seed_index=1, lmax=lmax)
        return m

    mix = model(S.SWM_MIXED)
    ref = model(S.SWM_FP64)
    h0, u0, tr0 = ref.field("h"), ref.field("u"), ref.field("tracer")
    mass0 = mix.total_mass()
    tm = timed(mix, args.warmup, args.steps)
    print(f"[bench] mixed {tm['wall'] * 1e3 / args.steps:.3f} ms/step", file=sys.stderr, flush=True)
    tf = timed(ref, args.warmup, args.steps)
    print(f"[bench] fp64  {tf['wall'] * 1e3 / args.steps:.3f} ms/step", file=sys.stderr, flush=True)
    value = cells * args.steps / tm["wall"]
    value64 = cells * args.steps / tf["wall"]
    # the mixed step on the FP64 step's solver schedule (same warm start)
    mrs = model(S.SWM_MIXED, reference_schedule=True)
    trs = timed(mrs, args.warmup, args.steps)
    mrs.close()
    value_rs = cells * args.steps / trs["wall"]
    print(f"[bench] mixed on the FP64 schedule {trs['wall'] * 1e3 / args.steps:.3f} ms/step", file=sys.stderr,
          flush=True)

    # the rest of one simulated day, timed per form; the error against FP64
    nerr = int(round(86400.0 / dt)) if args.error_steps is None else args.error_steps
    err, day = None, None
    done = args.warmup + args.steps
    if nerr > done:
        rest = {}
        for name, m in (("mixed", mix), ("fp64", ref)):
            m.synchronize()
            t1 = time.perf_counter()
            k = done
            while k < nerr:
                n = min(256, nerr - k)
                m.step(n)
                k += n
            m.synchronize()
            rest[name] = time.perf_counter() - t1
        day = {}
        for name, m, t in (("mixed", mix, tm), ("fp64", ref, tf)):
            wall = t["warm_wall"] + t["wall"] + rest[name]
            solves, vcyc, checks = m.solver_stats()
            day[name] = {"steps": nerr, "wall_s": wall, "cell_steps_per_s": cells * nerr / wall,
                         "ms_per_step": wall * 1e3 / nerr, "vcycles_per_step": vcyc / max(1, solves),
                         "checks_per_step": checks / max(1, solves)}
        day["mixed_over_fp64"] = day["mixed"]["cell_steps_per_s"] / day["fp64"]["cell_steps_per_s"]
        print(f"[bench] one day: mixed {day['mixed']['ms_per_step']:.3f} ms/step "
              f"({day['mixed']['vcycles_per_step']:.2f} cycles), fp64 {day['fp64']['ms_per_step']:.3f} ms/step "
              f"({day['fp64']['vcycles_per_step']:.2f} V-cycles)", file=sys.stderr, flush=True)
        # both forms solved to mean(r^2) <= 1e-24 from the same state: the
        # precision alone, and the converged solution against which the
        # solver-criterion error of both forms shows
        conv = model(S.SWM_FP64, tol=1e-24)
        mixt = model(S.SWM_MIXED, tol=1e-24)
        for m in (conv, mixt):
            k = 0
            while k < nerr:
                n = min(256, nerr - k)
                m.step(n)
                k += n
            m.synchronize()
        names = ("h", "eta", "div", "tracer")
        fm, fr, fc, ft = ({n: m.field(n) for n in names} for m in (mix, ref, conv, mixt))
        err = {"vs_fp64": rel_errors(grid, fr, fm), "vs_fp64_converged": rel_errors(grid, fc, fm),
               "fp64_vs_fp64_converged": rel_errors(grid, fc, fr),
               "precision_only": rel_errors(grid, fc, ft),
               "steps": nerr, "simulated_hours": nerr * dt / 3600.0,
               "mass_drift_mixed": abs(mix.total_mass() - mass0) / mass0,
               "h_target_1e-6": {"rel_l2_vs_fp64": None, "rel_linf_vs_fp64": None},
               "note": "relative L2 (area-weighted) and Linf over real cells; 'converged' is the FP64 form "
                       "solved to mean(r^2) <= 1e-24, the others use the reference criterion 1e-20; "
                       "'precision_only' is the mixed form against the FP64 form, both solved to 1e-24"}
        err["h_target_1e-6"] = {"rel_l2_vs_fp64": err["vs_fp64"]["h"]["rel_l2"],
                                "rel_linf_vs_fp64": err["vs_fp64"]["h"]["rel_linf"],
                                "rel_l2_precision_only": err["precision_only"]["h"]["rel_l2"],
                                "rel_linf_precision_only": err["precision_only"]["h"]["rel_linf"]}
        conv.close()
        mixt.close()
        print(f"[bench] error after {nerr} steps: h vs fp64 {err['vs_fp64']['h']}, precision only "
              f"{err['precision_only']['h']}, fp64 vs converged {err['fp64_vs_fp64_converged']['h']}",
              file=sys.stderr, flush=True)

    scan = None
    if not args.no_dt_scan:
        scan = dt_scan(S, grid, lmax, args.ntracer)
        print(f"[bench] dt scan {scan}", file=sys.stderr, flush=True)

    def kroof(t):
        if not t["kernel_ms"]:
            return None
        return t["kernel_bytes"] / (t["kernel_ms"] * 1e-3) / 1e9

    # HBM bytes per launch of the roofline kernels, measured with rocprofv3
    # FETCH_SIZE / WRITE_SIZE passes of this workload (profiles/traffic_g9_pmc.md)
    traffic = {}
    tpath = os.path.join(ROOT, "profiles", "traffic_g9_pmc.json")
    if g == 9 and os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("kernels", {})
    ach = kroof(tm)
    ach64 = kroof(tf)
    b_mix = step_bytes("mixed", args.ntracer, tm["vcycles_per_step"], tm["checks_per_step"])
    b_64 = step_bytes("fp64", args.ntracer, tf["vcycles_per_step"], tf["checks_per_step"])

    cpu = None
    if not args.no_cpu_baseline and args.cpu_steps > 0:
        try:
            ncores = min(16, len(os.sched_getaffinity(0)))
        except AttributeError:
            ncores = min(16, os.cpu_count() or 1)
        state = (h0, u0, tr0)
        v_all, t_all = cpu_baseline_run(grid, args.ntracer, dt, state, args.cpu_steps, ncores)
        v_one, t_one = cpu_baseline_run(grid, args.ntracer, dt, state, args.cpu_steps, 1)
        cpu = {"value": v_all, "unit": "cell-steps/s", "cores": ncores, "kind": "port",
               "sample": f"{args.cpu_steps} FP64 steps (the first ones) of the same g={g} workload from the same "
                         f"initial state, {t_all:.1f} s on {ncores} host threads (OpenMP over the 10 panels; "
                         f"global sums sequential); oracle/swm_oracle.c (gcc -O2), the C restatement of the "
                         f"reference step, bit-identical to the compiled reference at any thread count",
               "one_core": {"value": v_one, "cores": 1, "seconds": t_one}}
        print(f"[bench] cpu baseline {v_all:.3e} cell-steps/s on {ncores} threads, {v_one:.3e} on one",
              file=sys.stderr, flush=True)

    remap = None
    if not args.no_remap:
        remap = remap_bench(S, grid, dt, lmax, args.ntracer, args.warmup, args.steps)
        remap["fct_fp64_ms_per_step"] = tf["wall"] * 1e3 / args.steps
        remap["remap_over_fct_fp64_time"] = remap["ms_per_step"] / remap["fct_fp64_ms_per_step"]
        print(f"[bench] remap fp64 {remap['ms_per_step']:.3f} ms/step, kernels "
              + ", ".join(f"{k} {v['us']:.1f} us {v['gbs']:.0f} GB/s" for k, v in remap["kernels"].items()),
              file=sys.stderr, flush=True)

    lev8 = None
    if not args.no_level8:
        lev8 = level8_bench(S)
        print(f"[bench] level 8 FP64 day {lev8['ms_per_step']:.3f} ms/step, operators "
              + ", ".join(f"{k} {v['us']:.1f} us {v['gbs']:.0f} GB/s" for k, v in lev8["operators"].items()),
              file=sys.stderr, flush=True)

    adj = None
    if not args.no_adjoint:
        adj = adjoint_bench(S)
        print(f"[bench] adjoint {adj['adjoint_ms']:.0f} ms vs forward {adj['forward_ms']:.0f} ms "
              f"(ratio {adj['adj_over_fwd']:.2f}), dot {adj['dot_product_rel']:.1e}, "
              f"fd max {max(adj['fd_rel_errors']):.1e}; with the tracer {adj['with_tracer']['adjoint_ms']:.0f} ms "
              f"(ratio {adj['with_tracer']['adj_over_fwd']:.2f}), dot {adj['with_tracer']['dot_product_rel']:.1e}",
              file=sys.stderr, flush=True)

    ensemble = None
    if not args.no_ensemble:
        ensemble = ensemble_bench(S)
        print(f"[bench] ensemble {ensemble['member_steps_per_s']:.0f} member-steps/s "
              f"({ensemble['ms_per_ensemble_step']:.2f} ms per ensemble step, eager "
              f"{ensemble['eager_ms_per_ensemble_step']:.2f}), fixed cycles {ensemble['fixed_cycles']}, "
              f"above criterion {ensemble['solves_above_criterion']}, choice {ensemble['autotuner_choice']}",
              file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tm["wall"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 defect correction, FP64 sums)",
        "data": "synthetic (seeded harmonic fields, degree 1..%d, built on the device)" % lmax,
        "config": {
            "workload": f"icosahedral level {g} ({cells} cells), mixed-precision step, dt={dt} s, FCT mass + "
                        f"{args.ntracer} tracer, multigrid to mean(resid^2)<=1e-20",
            "g": g, "cells": cells, "dt": dt, "ntracer": args.ntracer,
            "radius": RADIUS, "grav": GRAV, "omega": OMEGA,
            "fp64_value": value64,
            "fp64_ms_per_step": tf["wall"] * 1e3 / args.steps,
            "mixed_over_fp64": value / value64,
            "vcycles_per_step": {"mixed": tm["vcycles_per_step"], "fp64": tf["vcycles_per_step"]},
            "same_warm_start": {"mixed_value": value_rs, "mixed_ms_per_step": trs["wall"] * 1e3 / args.steps,
                                "mixed_over_fp64": value_rs / value64,
                                "vcycles_per_step": trs["vcycles_per_step"],
                                "note": "the mixed step with the FP64 step's solver schedule: the last solution as "
                                        "the warm start and two cycles before the first convergence test"},
            "one_day": day,
            "mixed_error_vs_fp64": err,
            "dt_scan": scan,
            "remap": remap,
            "level8": lev8,
            "adjoint": adj,
            "ensemble": ensemble,
        },
        "roofline": {"bound": "hbm", "kernel": "k_relax_h<2> finest level (mixed path: FP32 fields, half-precision "
                               "weights)", "achieved": ach,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS if ach else None,
                     "traffic": traffic.get("k_relax_h<2>", {}).get("traffic_bytes"),
                     "traffic_note": "HBM bytes per two-field sweep, rocprofv3 FETCH_SIZE x 1.975 + WRITE_SIZE "
                                     "(scales from a plain copy kernel), profiles/traffic_g9_pmc.md",
                     "launches": tm["kernel_launches"], "mean_ms": tm["kernel_ms"],
                     "bytes_per_launch": tm["kernel_bytes"],
                     "fp64_kernel": {"kernel": "k_relax<double> finest level", "achieved": ach64,
                                     "frac": ach64 / HBM_PEAK_GBS if ach64 else None, "mean_ms": tf["kernel_ms"],
                                     "traffic": traffic.get("k_relax<double, 2, false>", {}).get("traffic_bytes"),
                                     "bytes_per_launch": tf["kernel_bytes"]}},
        "step_roofline": {"bytes_per_cell_step": {"mixed": b_mix, "fp64": b_64},
                          "achieved_gbs": {"mixed": value * b_mix / 1e9, "fp64": value64 * b_64 / 1e9},
                          "frac": {"mixed": value * b_mix / 1e9 / HBM_PEAK_GBS,
                                   "fp64": value64 * b_64 / 1e9 / HBM_PEAK_GBS}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
