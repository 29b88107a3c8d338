"""Benchmark of the MI355X shallow-water dynamical core.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--g 9] [--precision fp64]

One "step" is one full model time step (FCT mass and tracer transport,
vorticity/divergence ABM3 update, multigrid Poisson solve to the reference's
convergence criterion, velocity and kinetic energy, halo fills) over the whole
sphere.  The workload is the level-9 icosahedral grid (2,621,442 cells,
dt = 28.125 s) with seeded synthetic fields of harmonic degree 1..160, built on
the device; inputs are resident in HBM before the timed region starts.

Prints one JSON line: throughput in cell-steps/s, the roofline of the hottest
kernel (finest-level multigrid smoother, timed with HIP events on the model's
stream during the timed steps) and the CPU baseline (the oracle, a
single-threaded C restatement of the reference step, timed on one bounded
sample of the same workload).
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--g", type=int, default=9)
    ap.add_argument("--precision", choices=["fp64"], default="fp64")
    ap.add_argument("--ntracer", type=int, default=1)
    ap.add_argument("--lmax", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the panel-sharded multi-GPU design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import ctypes

    import swm_mi355x as S

    g = args.g
    dt = 450.0 * 2.0 ** (5 - g)
    lmax = args.lmax or {5: 20, 6: 40, 7: 80, 8: 80, 9: 160}.get(g, 20)
    radius, grav, omega = 6.371e6, 9.81, 7.292e-5
    t0 = time.time()
    grid = S.Grid(g, radius=radius, omega=omega)
    print(f"[bench] grid g={g} built in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    model = S.Model(grid, ntracer=args.ntracer, dt=dt, grav=grav)
    S._check(S.lib.swm_init_synthetic(model._h, 1, lmax, ctypes.c_double(40.0), ctypes.c_double(5000.0),
                                      ctypes.c_double(50.0), 4, ctypes.c_double(500.0e3)))
    h0 = model.field("h") if not args.no_cpu_baseline else None
    u0 = model.field("u") if not args.no_cpu_baseline else None
    t0tr = model.field("tracer") if not args.no_cpu_baseline else None
    mass0 = model.total_mass()

    model.step(args.warmup)
    model.synchronize()
    S._check(S.lib.swm_profile(model._h, 1))
    t1 = time.perf_counter()
    model.step(args.steps)
    model.synchronize()
    t2 = time.perf_counter()
    wall = t2 - t1
    step_ms = ctypes.c_double()
    kms = ctypes.c_double()
    kbytes = ctypes.c_double()
    nlaunch = ctypes.c_int()
    S._check(S.lib.swm_timing(model._h, ctypes.byref(step_ms), ctypes.byref(kms), ctypes.byref(kbytes),
                              ctypes.byref(nlaunch)))
    mass1 = model.total_mass()
    cells = grid.max_ig
    value = cells * args.steps / wall
    achieved = kbytes.value / (kms.value * 1e-3) / 1e9 if kms.value > 0 else None

    cpu = None
    if not args.no_cpu_baseline:
        from oracle.oracle import Oracle, grid_arrays_from

        orc = Oracle(grid_arrays_from(grid), args.ntracer, dt, grav)
        orc.init_state(h0, u0, t0tr)
        tc = time.perf_counter()
        orc.step(1)
        tcpu = time.perf_counter() - tc
        cpu = {"value": cells / tcpu, "unit": "cell-steps/s", "cores": 1, "kind": "port",
               "sample": f"1 step (the first, Euler-weighted) of the same g={g} workload from the same "
                         f"initial state, {tcpu:.1f} s on one core; oracle/swm_oracle.c, gcc -O2"}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded harmonic fields, degree 1..%d, built on the device)" % lmax,
        "config": {"workload": f"icosahedral level {g} ({cells} cells), {args.precision.upper()} plain step, "
                               f"dt={dt} s, FCT mass + {args.ntracer} tracer, multigrid to mean(resid^2)<=1e-20",
                   "g": g, "cells": cells, "dt": dt, "ntracer": args.ntracer, "precision": args.precision,
                   "radius": radius, "grav": grav, "omega": omega,
                   "vcycles_last_step": model.last_vcycles,
                   "device_ms_per_step": step_ms.value / args.steps,
                   "mass_drift": abs(mass1 - mass0) / mass0},
        "roofline": {"bound": "hbm", "kernel": "k_relax (finest level)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": None,
                     "launches": nlaunch.value, "mean_ms": kms.value, "bytes_per_launch": kbytes.value},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
