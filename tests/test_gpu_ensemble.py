"""Ensembles on the device: contexts of several members advanced in the same
launches (swm_create_members), and groups of them advanced together in
captured HIP graphs (swm_ensemble_*), with the autotuner of the launch choice.

Checks:
  - a member of an ensemble context is the same model as a single-member
    context from the same state: bit for bit, FP64 (the reference's
    criterion) and mixed (fixed schedule);
  - the perturbed synthetic states differ between members and agree with the
    single-member construction;
  - graph replays give, after every call, the same fields bit for bit as the
    groups stepped one launch at a time, for each launch choice (one stream
    or one per group, 1, 3 or 6 steps per graph, workgroup sizes), with two
    groups at different levels and substeps;
  - the autotuner returns the groups to the state and clock it started from;
  - the fixed schedule's residual record, and the refused calls.
"""
import numpy as np
import pytest

import swm_mi355x as S

pytestmark = pytest.mark.gpu

FIELDS = ("h", "eta", "div", "psi", "chi", "ke", "u", "tracer")


def _dt(g):
    return 200.0 * 2 ** (5 - g)


def _model(grid, precision, members=1, fixed=0):
    m = S.Model(grid, ntracer=1, dt=_dt(grid.g), grav=9.81, precision=precision, members=members)
    if fixed:
        m.set_fixed_cycles(fixed)
    return m


def _init(m, member, seed):
    m.select(member)
    m.init_synthetic_member(seed, 0.01, seed_index=0, lmax=12)


def _fields(m, member):
    m.select(member)
    return {n: m.field(n) for n in FIELDS}


@pytest.mark.parametrize("precision,fixed", [(S.SWM_FP64, 0), (S.SWM_MIXED, 2)])
def test_member_equals_single_context(precision, fixed):
    print(f"case members vs single precision={precision} fixed={fixed}", flush=True)
    grid = S.Grid(4)
    seeds = [0, 7, 0]
    ens = _model(grid, precision, members=3, fixed=fixed)
    for k, sd in enumerate(seeds):
        _init(ens, k, sd)
    ens.step(5)
    got = [_fields(ens, k) for k in range(3)]
    refs, cycles = [], set()
    for sd in seeds:
        one = _model(grid, precision, fixed=fixed)
        _init(one, 0, sd)
        one.step(5)
        refs.append(_fields(one, 0))
        cycles.add(one.solver_stats()[1])
    # the FP64 solve cycles every member while any misses the criterion: with
    # equal cycle counts the members are the single runs to the bit, else
    # they differ by what one more cycle changes
    exact = len(cycles) == 1
    for k in range(3):
        for n in FIELDS:
            if exact:
                assert np.array_equal(got[k][n], refs[k][n]), (k, n)
            else:
                d = np.abs(got[k][n] - refs[k][n]).max() / max(np.abs(refs[k][n]).max(), 1e-300)
                assert d <= 1e-9, (k, n, d)
    # different seeds give different members, equal seeds equal ones
    assert not np.array_equal(got[0]["eta"], got[1]["eta"])
    assert np.array_equal(got[0]["eta"], got[2]["eta"])
    ens.select(1)
    b1 = ens.budgets()
    ens.select(0)
    assert abs(b1[0] - ens.budgets()[0]) <= 1e-9 * b1[0]


def test_member_perturbation_size():
    print("case member perturbation size", flush=True)
    grid = S.Grid(4)
    m = _model(grid, S.SWM_FP64, members=2)
    m.select(0)
    m.init_synthetic(seed_index=0, lmax=12)
    m.select(1)
    m.init_synthetic_member(3, 0.01, seed_index=0, lmax=12)
    mask = grid.array("grid_center_mask") == 1.0
    z0 = m.select(0).field("relative")[mask]
    z1 = m.select(1).field("relative")[mask]
    # 1 % of the vorticity, then both velocity fields scaled to max |u| = 40
    rel = np.sqrt(np.mean((z1 - z0) ** 2) / np.mean(z0 ** 2))
    assert 0.002 < rel < 0.05, rel
    # zero perturbation is the plain construction
    m.select(1)
    m.init_synthetic_member(3, 0.0, seed_index=0, lmax=12)
    assert np.array_equal(m.select(1).field("u"), m.select(0).field("u"))


def _groups(fixed=2):
    ga, gb = S.Grid(3), S.Grid(4)
    a = _model(ga, S.SWM_MIXED, members=4, fixed=fixed)
    b = _model(gb, S.SWM_MIXED, members=2, fixed=fixed)
    for k in range(4):
        _init(a, k, k)
    for k in range(2):
        _init(b, k, 10 + k)
    return a, b


def _all(m):
    return [_fields(m, k) for k in range(m.members)]


def _same(x, y):
    return all(np.array_equal(p[n], q[n]) for p, q in zip(x, y) for n in FIELDS)


@pytest.mark.parametrize("choice", [dict(concurrent=0, steps_per_graph=1),
                                    dict(concurrent=1, steps_per_graph=3, blk_step=128, blk_mg=64),
                                    dict(concurrent=0, steps_per_graph=6, blk_step=64, blk_mg=128)])
def test_graph_replay_matches_eager(choice):
    print(f"case ensemble graph replay {choice}", flush=True)
    a1, b1 = _groups()
    a2, b2 = _groups()
    e = S.Ensemble([a2, b2], [1, 2])
    e.set_choice(**choice)
    # calls of 7, 6 and 5 ensemble steps: startup phases, whole graphs and
    # remainders, every replay compared with the groups stepped eagerly
    for n in (7, 6, 5):
        a1.step(n)
        b1.step(2 * n)
        e.step(n)
        assert _same(_all(a1), _all(a2)), n
        assert _same(_all(b1), _all(b2)), n
        assert a2.step_count == a1.step_count and b2.step_count == b1.step_count
        assert a2.time_seconds == a1.time_seconds and b2.time_seconds == b1.time_seconds
    assert e.choice()["ms_per_step"] > 0.0


def test_tuner_restores_state_and_clock():
    print("case ensemble tuner", flush=True)
    a1, b1 = _groups()
    a2, b2 = _groups()
    e = S.Ensemble([a2, b2], [1, 2])
    e.step(4)
    a1.step(4)
    b1.step(8)
    trials = e.tune(rounds=1)
    assert len(trials) >= 6 and all(t["ms_per_step"] > 0 for t in trials)
    best = e.choice()
    assert best["blk_step"] in (64, 128, 256) and best["blk_mg"] in (64, 128, 256)
    assert a2.step_count == 4 and b2.step_count == 8
    assert _same(_all(a1), _all(a2)) and _same(_all(b1), _all(b2))
    e.step(5)
    a1.step(5)
    b1.step(10)
    assert _same(_all(a1), _all(a2)) and _same(_all(b1), _all(b2))


def test_fixed_schedule_track_and_errors():
    print("case fixed schedule track and errors", flush=True)
    grid = S.Grid(4)
    m = _model(grid, S.SWM_MIXED, members=2, fixed=3)
    for k in range(2):
        _init(m, k, k)
    m.step(6)
    for k in range(2):
        m.select(k)
        r0, r1, above = m.solver_track()
        assert 0.0 < r0 <= 1e-20 and 0.0 < r1 <= 1e-20, (r0, r1)
        assert above == 0
    # the schedule needs the absolute criterion; graphs need the mixed path
    f = _model(grid, S.SWM_FP64, members=2)
    f.set_solver(tol=1e-12, relative=True)
    with pytest.raises(S.SwmError):
        f.set_fixed_cycles(2)
    f.set_solver()
    f.set_fixed_cycles(2)
    for k in range(2):
        _init(f, k, k)
    with pytest.raises(S.SwmError):
        S.Ensemble([f], [1]).step(1)
    with pytest.raises(S.SwmError):
        m.set_solver(tol=1e-10, relative=True)
    with pytest.raises(S.SwmError):
        m.select(2)
    # every member before the first step
    p = _model(grid, S.SWM_MIXED, members=2, fixed=2)
    _init(p, 0, 0)
    with pytest.raises(S.SwmError):
        p.step(1)
    # graphs need the fixed schedule; the linearised model one member
    q = _model(grid, S.SWM_MIXED, members=2)
    for k in range(2):
        _init(q, k, k)
    e = S.Ensemble([q], [1])
    with pytest.raises(S.SwmError):
        e.step(1)
    c = S.Model(grid, ntracer=1, dt=_dt(4), grav=9.81, advection="centered", members=2)
    for k in range(2):
        _init(c, k, k)
    with pytest.raises(S.SwmError):
        c.adj_forward(2, 1)
