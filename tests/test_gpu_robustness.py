"""Failure detection and debugging aids on the device.

  - a NaN in the state is reported by the step (SWM_ERR_NAN, from the
    residual norms every solve reads back) in both precisions;
  - the block-size self-test of the kernels' implicit arguments passes;
  - initialising one member of an ensemble context after steps lets the
    other members continue from their latest state (forward Euler restart),
    bit for bit as a single context restarted with swm_restart_clock;
  - the halo shadow check finds no difference between the device's halo
    fills and the host's, and changes no result.
"""
import numpy as np
import pytest

import swm_mi355x as S

pytestmark = pytest.mark.gpu

FIELDS = ("h", "eta", "div", "psi", "chi", "ke", "u", "tracer")


def _fields(m):
    return {n: m.field(n) for n in FIELDS}


@pytest.mark.parametrize("precision", [S.SWM_FP64, S.SWM_MIXED])
def test_nan_in_state_is_reported(precision):
    print(f"case NaN detection precision={precision}", flush=True)
    grid = S.Grid(4)
    m = S.Model(grid, ntracer=1, dt=400.0, grav=9.81, precision=precision)
    m.init_synthetic(lmax=8)
    h, u, tr = m.field("h"), m.field("u"), m.field("tracer")
    h[5, 7, 3] = np.nan
    m.init_state(h, u, tr)
    with pytest.raises(S.SwmError) as e:
        m.step(2)
    assert e.value.status == S.SWM_ERR_NAN, e.value


def test_block_size_self_test():
    assert S.lib.swm_self_test(0) == 0
    assert S.lib.swm_self_test(4096) == S.SWM_ERR_NO_DEVICE


def test_member_reinit_continues_the_others_from_their_latest_state():
    print("case member re-initialisation after an odd step count", flush=True)
    grid = S.Grid(4)

    def model(members):
        m = S.Model(grid, ntracer=1, dt=400.0, grav=9.81, precision=S.SWM_MIXED, members=members)
        m.set_fixed_cycles(2, 4)
        return m

    a = model(2)
    for k, seed in ((0, 0), (1, 7)):
        a.select(k)
        a.init_synthetic_member(seed, 0.01, lmax=8)
    a.step(3)
    a.select(1)
    a.init_synthetic_member(9, 0.01, lmax=8)
    a.step(4)
    b = model(1)
    b.init_synthetic_member(0, 0.01, lmax=8)
    b.step(3)
    b.restart_clock()
    b.step(4)
    fa, fb = _fields(a.select(0)), _fields(b)
    for n in FIELDS:
        assert np.array_equal(fa[n], fb[n]), n
    # the re-initialised member is a fresh run of its own state
    c = model(1)
    c.init_synthetic_member(9, 0.01, lmax=8)
    c.step(4)
    fa1, fc = _fields(a.select(1)), _fields(c)
    for n in FIELDS:
        assert np.array_equal(fa1[n], fc[n]), n


@pytest.mark.parametrize("precision", [S.SWM_FP64, S.SWM_MIXED])
def test_halo_shadow_check(precision):
    print(f"case halo shadow precision={precision}", flush=True)
    grid = S.Grid(4)
    out = []
    for shadow in (False, True):
        m = S.Model(grid, ntracer=2, dt=400.0, grav=9.81, precision=precision)
        m.set_halo_shadow(shadow)
        m.init_synthetic(lmax=8)
        m.step(4)
        out.append(_fields(m))
        if shadow:
            checks, bad = m.halo_shadow_stats()
            assert checks > 50 and bad == 0, (checks, bad)
    for n in FIELDS:
        assert np.array_equal(out[0][n], out[1][n]), n
    # not inside captured graphs
    e = S.Model(grid, ntracer=1, dt=400.0, grav=9.81, precision=S.SWM_MIXED)
    e.set_fixed_cycles(1)
    e.init_synthetic(lmax=8)
    e.set_halo_shadow(True)
    ens = S.Ensemble([e], [1])
    with pytest.raises(S.SwmError):
        ens.step(1)
