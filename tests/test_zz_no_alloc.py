"""The allocation-free global sum (swm_global_real_sum_noalloc, device arrays,
block partials in a file-scope device array) against the context entry point
it varies (swm_global_real_sum_device, partials in the context's buffers):
the same kernels, so the same sums bit for bit, on two calls in a row with
the second smaller than the first (the array keeps the first call's
partials; each call writes all of them), and the size limit refused.  The
test's device arrays come from the HIP runtime the library links."""
import ctypes
import math
import os

import numpy as np
import pytest

import swm_mi355x as S

pytestmark = pytest.mark.gpu


class _Hip:
    H2D, D2H = 1, 2

    def __init__(self):
        # the runtime the library links (torch, when imported, brings its own
        # copy under the same soname, which has no device of this process)
        path = "/opt/rocm/lib/libamdhip64.so.7"
        self.lib = ctypes.CDLL(path if os.path.exists(path) else "libamdhip64.so.7")
        self.bufs = []

    def _ok(self, e):
        assert e == 0, f"HIP error {e}"

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        self._ok(self.lib.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)))
        self.bufs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.alloc(max(a.nbytes, 8))
        self._ok(self.lib.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), self.H2D))
        return p

    def get(self, p, n):
        out = np.zeros(n)
        self._ok(self.lib.hipDeviceSynchronize())
        self._ok(self.lib.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(8 * n), self.D2H))
        return out

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def test_noalloc_sum_matches_context_sum():
    print("case allocation-free global sum", flush=True)
    grid = S.Grid(3)
    m = S.Model(grid, ntracer=1, dt=800.0, grav=S.REF_GRAV)
    hip = _Hip()
    rng = np.random.default_rng(9)
    try:
        for n in (1_000_003, 5_001):
            x = rng.standard_normal(n) * 10.0 ** rng.integers(-6, 6, n)
            mask = (rng.random(n) < 0.6).astype(np.float64)
            dx, dm = hip.put(x), hip.put(mask)
            out_ctx, out_na = hip.put(np.zeros(2)), hip.put(np.zeros(2))
            second = lambda p: ctypes.c_void_p(p.value + 8)
            S._check(S.lib.swm_global_real_sum_device(m._h, dx, dm, n, out_ctx))
            S._check(S.lib.swm_global_real_sum_device(m._h, dx, None, n, second(out_ctx)))
            m.synchronize()
            S._check(S.lib.swm_global_real_sum_noalloc(dx, dm, n, out_na, 0, None))
            S._check(S.lib.swm_global_real_sum_noalloc(dx, None, n, second(out_na), 0, None))
            a, b = hip.get(out_ctx, 2), hip.get(out_na, 2)
            assert a[0] == b[0] and a[1] == b[1], (n, a, b)
            assert abs(b[0] - math.fsum(x * mask)) <= 4e-16 * np.sum(np.abs(x * mask))
            assert abs(b[1] - math.fsum(x)) <= 4e-16 * np.sum(np.abs(x))
        # beyond the maximum, or no such device: refused before any launch
        with pytest.raises(S.SwmError):
            S._check(S.lib.swm_global_real_sum_noalloc(dx, None, (1 << 30) + 1, out_na, 0, None))
        with pytest.raises(S.SwmError):
            S._check(S.lib.swm_global_real_sum_noalloc(dx, None, 10, out_na, 4096, None))
    finally:
        hip.free()
