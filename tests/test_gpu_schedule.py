"""Schedule options of the step that leave the arithmetic alone or only
change how the mixed solve converges:

  - stream overlap (swm_set_overlap): the tracer transport on a second
    stream, started at once or by the first V-cycle, gives the same fields
    bit for bit, FP64 and mixed;
  - the multigrid kernel fusions (swm_set_fusion) give the same fields bit
    for bit as the unfused cycle, FP64 and mixed, for two finest-level sweep
    counts, over levels 6..4 with their own sweeps;
  - the producer-side halo writes of the step (SWM_FUSE_HALO_PUSH) give the
    same fields, halo slots included, bit for bit as the step's halo fills,
    FP64 and mixed, one and two tracers, with and without the tracer
    stream, for every member of a two-member context;
  - the velocity kernel that a one-member mixed step queues at each test
    of its solve (run only where the device's copy of the test passes)
    gives the same fields, and the same cycles, as a two-member context
    with both members in the same state, which launches it after the test;
  - the mixed solve's coarse correction (swm_set_mixed_coarse): the dense
    solves of level 4 and of level 3 and the reference's V-cycle below level
    4 all meet the criterion, with errors against the FP64 form within 3x
    of each other plus 1e-7, and the dense solves take no more cycles;
  - the mixed solve's finest-level sweep count (swm_set_mixed_smoother):
    every count meets the convergence criterion (the V-cycle only
    preconditions the FP64 defect correction), and its error against the
    FP64 form is within 3x of the two-sweep run's plus 1e-7 (thickness,
    vorticity, divergence, relative L2 after 8 steps).
"""
import numpy as np
import pytest

import swm_mi355x as S

pytestmark = pytest.mark.gpu

FIELDS = ("h", "eta", "div", "psi", "chi", "u", "tracer")


def _model(grid, precision, **kw):
    m = S.Model(grid, ntracer=2, dt=200.0 * 2 ** (5 - grid.g), grav=9.81, precision=precision)
    if "overlap" in kw:
        m.set_overlap(kw["overlap"])
    if "sweeps" in kw:
        m.set_mixed_smoother(kw["sweeps"])
    if "fusion" in kw:
        m.set_fusion(kw["fusion"])
    if "coarse_direct" in kw:
        m.set_mixed_coarse(kw["coarse_direct"], kw.get("coarse_level", 4))
    m.init_synthetic(seed_index=2, lmax=12)
    return m


@pytest.mark.parametrize("precision,mode", [(S.SWM_FP64, 1), (S.SWM_MIXED, 1), (S.SWM_FP64, 3), (S.SWM_MIXED, 2),
                                            (S.SWM_MIXED, 3)])
def test_overlap_gives_the_same_fields(precision, mode):
    print(f"case overlap precision={precision} mode={mode}", flush=True)
    grid = S.Grid(6)
    a = _model(grid, precision, overlap=0)
    b = _model(grid, precision, overlap=mode)
    a.step(7)
    b.step(7)
    for n in FIELDS:
        assert np.array_equal(a.field(n), b.field(n)), n


@pytest.mark.parametrize("precision,sweeps", [(S.SWM_FP64, 2), (S.SWM_MIXED, 2), (S.SWM_MIXED, 1)])
def test_fusion_gives_the_same_fields(precision, sweeps):
    print(f"case multigrid fusion precision={precision} sweeps={sweeps}", flush=True)
    grid = S.Grid(6)
    kw = {"sweeps": sweeps} if precision == S.SWM_MIXED else {}
    a = _model(grid, precision, fusion=0, **kw)
    b = _model(grid, precision, **kw)
    a.step(5)
    b.step(5)
    assert a.solver_stats() == b.solver_stats()
    for n in FIELDS:
        assert np.array_equal(a.field(n), b.field(n)), n
    with pytest.raises(S.SwmError):
        b.set_fusion(S.FUSE_ALL + 8)


@pytest.mark.parametrize("precision,ntracer,overlap", [(S.SWM_FP64, 2, 3), (S.SWM_MIXED, 2, 3), (S.SWM_MIXED, 1, 0),
                                                     (S.SWM_FP64, 1, 1)])
def test_halo_push_gives_the_same_fields(precision, ntracer, overlap):
    print(f"case halo push precision={precision} ntracer={ntracer} overlap={overlap}", flush=True)
    grid = S.Grid(6)
    dt = 200.0 * 2 ** (5 - grid.g)
    ms = []
    for flags in (S.FUSE_INJECT_SWEEP, S.FUSE_ALL):
        m = S.Model(grid, ntracer=ntracer, dt=dt, grav=9.81, precision=precision, members=2)
        m.set_overlap(overlap)
        m.set_fusion(flags)
        for z in range(2):
            m.select(z).init_synthetic_member(100 + z, lmax=12, seed_index=2)
        m.step(6)
        ms.append(m)
    a, b = ms
    for z in range(2):
        a.select(z)
        b.select(z)
        for n in FIELDS + ("ke", "relative"):
            assert np.array_equal(a.field(n), b.field(n)), (z, n)


@pytest.mark.parametrize("sweeps", [2, 1])
def test_early_velocity_gives_the_same_fields(sweeps):
    # one member: the step's velocity kernel is queued at each test of the
    # solve and runs only where the device's copy of the test passes; two
    # members (no early launch) with identical states take the same cycles
    print(f"case early velocity sweeps={sweeps}", flush=True)
    grid = S.Grid(6)
    dt = 200.0 * 2 ** (5 - grid.g)
    ms = []
    for members in (1, 2):
        m = S.Model(grid, ntracer=1, dt=dt, grav=9.81, precision=S.SWM_MIXED, members=members)
        m.set_mixed_smoother(sweeps)
        for z in range(members):
            m.select(z).init_synthetic_member(7, lmax=12, seed_index=2)
        m.step(8)
        ms.append(m)
    one, two = ms
    cycles = one.solver_stats()
    assert cycles == two.solver_stats(), (cycles, two.solver_stats())
    for z in range(2):
        two.select(z)
        for n in FIELDS + ("ke", "relative"):
            assert np.array_equal(one.field(n), two.field(n)), (z, n)


def test_mixed_coarse_solve_meets_the_criterion():
    print("case mixed coarse solve dense vs V-cycle", flush=True)
    grid = S.Grid(6)
    fp = _model(grid, S.SWM_FP64)
    vc = _model(grid, S.SWM_MIXED, coarse_direct=False)
    de = _model(grid, S.SWM_MIXED, coarse_direct=True, coarse_level=3)
    d4 = _model(grid, S.SWM_MIXED, coarse_direct=True, coarse_level=4)
    for _ in range(8):
        for m in (fp, vc, de, d4):
            m.step(1)
        for m in (vc, de, d4):
            r, _ = m.last_residual()
            assert max(r) <= 1e-20, r
    assert de.solver_stats()[1] <= vc.solver_stats()[1], (de.solver_stats(), vc.solver_stats())
    assert d4.solver_stats()[1] <= vc.solver_stats()[1], (d4.solver_stats(), vc.solver_stats())
    mask = grid.array("grid_center_mask") == 1.0
    area = grid.array("area")[mask]

    def err(a, n):
        x, y = fp.field(n)[mask], a.field(n)[mask]
        return np.sqrt(np.sum(area * (y - x) ** 2) / np.sum(area * x * x))

    for n in ("h", "eta", "div"):
        ev = err(vc, n)
        for ed in (err(de, n), err(d4, n)):
            assert ed <= 3.0 * ev + 1e-7 and ev <= 3.0 * ed + 1e-7, (n, ev, ed)
    with pytest.raises(S.SwmError):
        fp.set_mixed_coarse(True)


@pytest.mark.parametrize("sweeps", [1, 3, 4])
def test_mixed_sweep_counts_meet_the_criterion(sweeps):
    print(f"case mixed finest sweeps={sweeps}", flush=True)
    grid = S.Grid(5)
    fp = _model(grid, S.SWM_FP64)
    ref = _model(grid, S.SWM_MIXED)
    m = _model(grid, S.SWM_MIXED, sweeps=sweeps)
    for _ in range(8):
        fp.step(1)
        ref.step(1)
        m.step(1)
        r, _ = m.last_residual()
        assert max(r) <= 1e-20, r
    # both solves stop anywhere below mean(r^2) = 1e-20, so the fields differ
    # by what the criterion leaves free; against the FP64 form, the sweep
    # count must not make the mixed step less accurate than two sweeps do
    mask = grid.array("grid_center_mask") == 1.0
    area = grid.array("area")[mask]

    def err(a, n):
        x, y = fp.field(n)[mask], a.field(n)[mask]
        return np.sqrt(np.sum(area * (y - x) ** 2) / np.sum(area * x * x))

    # (plus 1e-7, the precision bar of the mixed step: with the quadratic
    # warm start the two-sweep run's error can itself be far below it)
    for n in ("h", "eta", "div"):
        e2, es = err(ref, n), err(m, n)
        assert es <= 3.0 * e2 + 1e-7, (n, e2, es)
    with pytest.raises(S.SwmError):
        m.set_mixed_smoother(5)
