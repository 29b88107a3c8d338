"""TEST INFRASTRUCTURE ONLY: ctypes binding of the oracle (swm_oracle.c).

The oracle is the CPU FP64 restatement of the reference step used to check
the device path and timed as the CPU baseline; it is never part of the
shipped library.  Build with ``make -C oracle`` (writes oracle/liboracle.so).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [vp]
    lib.orc_destroy.argtypes = [vp]
    lib.orc_init.argtypes = [vp, vp, vp, vp]
    lib.orc_step.restype = ctypes.c_int
    lib.orc_step.argtypes = [vp, ctypes.c_int]
    lib.orc_get.restype = ctypes.c_int
    lib.orc_get.argtypes = [vp, ctypes.c_char_p, vp]
    lib.orc_last_vcycles.restype = ctypes.c_int
    lib.orc_last_vcycles.argtypes = [vp]
    for name in ("orc_div_of_an_edge_flux", "orc_laplacian", "orc_curl_of_a_vector", "orc_div_of_a_vector",
                 "orc_center_ke", "orc_average_to_corners"):
        getattr(lib, name).argtypes = [vp, vp, vp]
    lib.orc_velocity.argtypes = [vp, vp, vp, vp]
    lib.orc_centered_mass_flux.argtypes = [vp, vp, vp, vp]
    lib.orc_upstream_mass_flux.argtypes = [vp, vp, vp, vp]
    lib.orc_vorticity_flux.argtypes = [vp, vp, vp, vp, vp]
    lib.orc_divergence_curl.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.orc_multigrid.argtypes = [vp, ctypes.c_int, vp, vp]
    lib.orc_elliptic.restype = ctypes.c_int
    lib.orc_elliptic.argtypes = [vp, ctypes.c_int, vp, vp]
    lib.orc_wrap.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
    lib.orc_budgets.argtypes = [vp, vp]
    lib.orc_set_sum_mode.argtypes = [vp, ctypes.c_int]
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    lib.orc_set_mass_fct.argtypes = [vp] * 10
    lib.orc_set_tracer_fct.argtypes = [vp] * 8
    lib.orc_set_remap.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int]
    lib.orc_set_remap_limiter.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.orc_set_diss.argtypes = [vp] * 5 + [ctypes.c_double]
    lib.orc_dissipation.restype = ctypes.c_int
    lib.orc_dissipation.argtypes = [vp, ctypes.c_int, ctypes.c_double, vp, vp]
    lib.orc_remap_mass_and_tracers.restype = ctypes.c_int
    lib.orc_remap_mass_and_tracers.argtypes = [vp] * 5
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


REMAP_FIELDS = ("c", "w", "trans_to_x1", "trans_x1_to_x4", "trans_x4_to_x3", "cl_x4", "cr_x4", "tl_x4", "tr_x4",
                "bl_x4", "br_x4", "p_x4", "c_x3", "area_inv_x3", "xav", "yav", "xxav", "xyav", "yyav", "gwgt_x3",
                "garea_inv")


class _OrcGrid(ctypes.Structure):
    _fields_ = [("g", ctypes.c_int), ("ntracer", ctypes.c_int), ("dt", ctypes.c_double), ("grav", ctypes.c_double),
                ("scheme", ctypes.c_int)] + [
        (n, ctypes.c_void_p) for n in ("area", "area_inv", "corner_area_inv", "sub_area", "dc", "normal", "tangent",
                                       "v_weights", "s_weights", "l_weights", "f", "hs", "center_mask")] + [
        ("nface", ctypes.c_int), ("nvrtx", ctypes.c_int), ("nedge", ctypes.c_int)] + [
        (n, ctypes.c_void_p) for n in ("face_dst", "face_src", "vrtx_dst", "vrtx_src", "edge_dst", "edge_src")] + [
        ("lap_w", ctypes.c_void_p), ("tot_area", ctypes.c_double)]


class _OrcRemap(ctypes.Structure):
    _fields_ = [("radius", ctypes.c_double)] + [
        (n, ctypes.c_void_p) for n in REMAP_FIELDS]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


GRID_FIELDS = ["area", "area_inv", "corner_area_inv", "sub_area", "dc", "normal", "tangent", "v_weights",
               "s_weights", "l_weights", "f", "grid_center_mask", "face_dst", "face_src", "vrtx_dst", "vrtx_src",
               "edge_dst", "edge_src"]


def grid_arrays_from(grid):
    """Set-up arrays of a swm_mi355x.Grid (or any object with .array and .g)."""
    d = {n: grid.array(n) for n in GRID_FIELDS}
    for l in range(1, grid.g + 1):
        d[f"lap_w{l}"] = grid.array(f"lap_w{l}")
    d["g"] = grid.g
    d["tot_area"] = grid.array("tot_area")
    return d


SCHEMES = {"fct": 0, "centered": 1, "upwind": 2, "remap": 3}


def remap_arrays_from(grid):
    """The remapping quantities (init_remap) of a swm_mi355x.Grid, with the
    corners, edge midpoints and corner frames they use, for scheme "remap"."""
    d = {"radius": grid.radius, "c": grid.array("c"), "w": grid.array("w"), "trans_to_x1": grid.array("trans_to_X1")}
    for n in ("cl_x4", "cr_x4", "tl_x4", "tr_x4", "bl_x4", "br_x4", "p_x4", "c_x3", "area_inv_x3", "xav", "yav", "xxav",
              "xyav", "yyav", "gwgt_x3", "garea_inv"):
        d[n] = grid.array("remap_" + n)
    d["trans_x1_to_x4"] = grid.array("remap_trans_X1_to_X4")
    d["trans_x4_to_x3"] = grid.array("remap_trans_X4_to_X3")
    return d


def set_threads(n):
    """Host threads of the oracle's panel-parallel loops (results do not
    depend on the count)."""
    lib().orc_set_threads(int(n))


class Oracle:
    """One FP64 model state on the CPU; arrays as in the reference layout.
    scheme: the advection of step_dt, "fct" (l_fct), "centered" (l_centered),
    "upwind" (l_upwind) or "remap" (l_remap; remap = remap_arrays_from(grid),
    limiter = the tracer limiter's (l_limit_tracers_min, _max)).  sums: the elliptic solve's global sums in the
    reference's memory order ("reference") or compensated ("compensated")."""

    def __init__(self, arrays, ntracer, dt, grav, hs=None, scheme="fct", sums="reference", remap=None,
                 limiter=(1, 1)):
        L = lib()
        self.g = int(arrays["g"])
        self.iim = self.jjm = 2 + 2 ** self.g
        self.ntracer = ntracer
        ns = self.iim * self.jjm * 10
        keep = {}

        def f64(name, a):
            keep[name] = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, order="F"))
            return _p(keep[name])

        def i32(name, a):
            keep[name] = np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1))
            return _p(keep[name])

        gr = _OrcGrid()
        gr.g, gr.ntracer, gr.dt, gr.grav = self.g, ntracer, dt, grav
        gr.scheme = SCHEMES[scheme]
        for n in ("area", "area_inv", "corner_area_inv", "sub_area", "dc", "normal", "tangent", "v_weights",
                  "s_weights", "l_weights", "f"):
            setattr(gr, n, f64(n, arrays[n]))
        gr.center_mask = f64("mask", arrays["grid_center_mask"])
        gr.hs = f64("hs", np.zeros(ns) if hs is None else hs)
        for n in ("face", "vrtx", "edge"):
            setattr(gr, "n" + n, len(arrays[n + "_dst"]))
            setattr(gr, n + "_dst", i32(n + "_dst", arrays[n + "_dst"]))
            setattr(gr, n + "_src", i32(n + "_src", arrays[n + "_src"]))
        ptrs = (ctypes.c_void_p * self.g)()
        for l in range(1, self.g + 1):
            ptrs[l - 1] = f64(f"lap_w{l}", arrays[f"lap_w{l}"])
        gr.lap_w = ctypes.cast(ptrs, ctypes.c_void_p)
        gr.tot_area = float(arrays["tot_area"])
        self._keep, self._ptrs = keep, ptrs
        self._h = L.orc_create(ctypes.byref(gr))
        L.orc_set_sum_mode(self._h, {"reference": 0, "compensated": 1}[sums])
        if remap is not None:
            rm = _OrcRemap()
            rm.radius = float(remap["radius"])
            for n in REMAP_FIELDS:
                setattr(rm, n, f64("remap_" + n, remap[n]))
            L.orc_set_remap(self._h, ctypes.byref(rm), int(limiter[0]), int(limiter[1]))

    def init_state(self, h, u, tracer):
        h = np.asfortranarray(h, dtype=np.float64)
        u = np.asfortranarray(u, dtype=np.float64)
        t = np.asfortranarray(tracer, dtype=np.float64)
        lib().orc_init(self._h, _p(h), _p(u), _p(t))

    def step(self, n=1):
        st = lib().orc_step(self._h, n)
        if st == 1:
            raise RuntimeError("oracle: elliptic solve did not converge")
        if st == 2:
            raise RuntimeError("oracle: remapping met a negative thickness (or has no remapping quantities)")

    def field(self, name):
        s = (self.iim, self.jjm, 10)
        shape = {"u": (2, 2) + s, "tracer": (self.iim, self.jjm, self.ntracer, 10)}.get(name, s)
        out = np.zeros(shape, order="F")
        if lib().orc_get(self._h, name.encode(), _p(out)) != 0:
            raise KeyError(name)
        return out

    @property
    def last_vcycles(self):
        return lib().orc_last_vcycles(self._h)

    def op(self, name, *args, out_shape):
        out = np.zeros(out_shape, order="F")
        arrs = [np.asfortranarray(a, dtype=np.float64) for a in args]
        getattr(lib(), name)(self._h, *[_p(a) for a in arrs], _p(out))
        return out

    def wrap(self, kind, x, nspace=1):
        """Halo fill of one layer in place: kind "face" (iim,jjm,10), "vrtx"
        (nspace,2,iim,jjm,10), "edge" (3,iim,jjm,10)."""
        assert x.dtype == np.float64 and x.flags.f_contiguous
        lib().orc_wrap(self._h, {"face": 0, "vrtx": 1, "edge": 2}[kind], nspace, _p(x))
        return x

    def budgets(self):
        """compute_budgets global averages (budgets.F:31-247): mass, ke, pe,
        tf, te, potential enstrophy, eta, div, then one mass-weighted
        variance per tracer."""
        out = np.zeros(8 + self.ntracer)
        lib().orc_budgets(self._h, _p(out))
        return out

    def multigrid(self, beta, alpha):
        nvar = beta.shape[2]
        b = np.ascontiguousarray(np.moveaxis(np.asarray(beta, dtype=np.float64), 2, 0))  # (nvar, iim, jjm, 10)
        a = np.ascontiguousarray(np.moveaxis(np.asarray(alpha, dtype=np.float64), 2, 0))
        bf = np.stack([x.reshape(-1, order="F") for x in b])
        af = np.stack([x.reshape(-1, order="F") for x in a])
        lib().orc_multigrid(self._h, nvar, _p(bf), _p(af))
        return np.stack([x.reshape((self.iim, self.jjm, 10), order="F") for x in af], axis=2)

    def elliptic(self, rhs, xout):
        nvar = rhs.shape[2]
        rf = np.stack([np.asarray(rhs[:, :, v, :], dtype=np.float64).reshape(-1, order="F") for v in range(nvar)])
        xf = np.stack([np.asarray(xout[:, :, v, :], dtype=np.float64).reshape(-1, order="F") for v in range(nvar)])
        st = lib().orc_elliptic(self._h, nvar, _p(rf), _p(xf))
        shp = (self.iim, self.jjm, 10)
        r = np.stack([x.reshape(shp, order="F") for x in rf], axis=2)
        x = np.stack([x.reshape(shp, order="F") for x in xf], axis=2)
        return st, r, x

    def set_mass_fct(self, tend_index, tend_weights, mss_n3, mss_c, u, msh_f):
        """set_mass_fct (mass_fct.F:23-539): msh_f (3, iim, jjm, 3 slots, 10)
        in the reference layout, updated in its slot tend_index[0]; returns
        (new mass, low-order mass, total mass flux (3, iim, jjm, 10))."""
        s = (self.iim, self.jjm, 10)
        ti = np.ascontiguousarray(tend_index, dtype=np.int32)
        tw = np.ascontiguousarray(tend_weights, dtype=np.float64)
        hist = np.ascontiguousarray(np.stack([np.asarray(msh_f)[:, :, :, t, :].reshape(-1, order="F")
                                              for t in range(3)]))
        args = [np.asfortranarray(a, dtype=np.float64) for a in (mss_n3, mss_c, u)]
        low, new = np.zeros(s, order="F"), np.zeros(s, order="F")
        flux = np.zeros((3,) + s, order="F")
        lib().orc_set_mass_fct(self._h, _p(ti), _p(tw), *[_p(a) for a in args], _p(hist), _p(low), _p(flux), _p(new))
        for t in range(3):
            msh_f[:, :, :, t, :] = hist[t].reshape((3,) + s, order="F")
        return new, low, flux

    def set_tracer_fct(self, trc, trc_high, mss, mss_low, mss_horz, mflux):
        """set_tracer_fct (tracer_fct.F:22-670): the flux-corrected tracer."""
        args = [np.asfortranarray(a, dtype=np.float64) for a in (trc, trc_high, mss, mss_low, mss_horz, mflux)]
        out = np.zeros((self.iim, self.jjm, 10), order="F")
        lib().orc_set_tracer_fct(self._h, *[_p(a) for a in args], _p(out))
        return out

    def set_diss(self, grid):
        """init_diss_op from a swm_mi355x.Grid's centres, corners, edge
        midpoints and corner frames."""
        arrs = [np.asfortranarray(grid.array(n), dtype=np.float64) for n in ("p", "c", "w", "trans_to_X2")]
        lib().orc_set_diss(self._h, *[_p(a) for a in arrs], float(grid.radius))

    def dissipation(self, norder, visc, xin):
        """dissipation (swm_dissipation_operator.F:212-443) of corner vectors
        (2, 2, iim, jjm, 10)."""
        x = np.asfortranarray(xin, dtype=np.float64)
        out = np.zeros_like(x, order="F")
        if lib().orc_dissipation(self._h, int(norder), float(visc), _p(x), _p(out)) != 0:
            raise RuntimeError("oracle: set_diss first")
        return out

    def set_remap_limiter(self, limit_min, limit_max):
        lib().orc_set_remap_limiter(self._h, int(limit_min), int(limit_max))

    def remap_mass_and_tracers(self, mass, trace, uc):
        """remap_mass_and_tracers (remap.F:1072-1604) with the model's dt:
        returns (mass, trace, edgeflux (3, iim, jjm, 10))."""
        m = np.array(mass, dtype=np.float64, order="F")
        t = np.array(trace, dtype=np.float64, order="F")
        u = np.asfortranarray(uc, dtype=np.float64)
        ef = np.zeros((3, self.iim, self.jjm, 10), order="F")
        if lib().orc_remap_mass_and_tracers(self._h, _p(m), _p(t), _p(u), _p(ef)) != 0:
            raise RuntimeError("oracle: remapping met a negative mass")
        return m, t, ef

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
