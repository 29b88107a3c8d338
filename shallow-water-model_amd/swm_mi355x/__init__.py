"""Python binding of the MI355X shallow-water dynamical core (ctypes over the
C ABI declared in include/swm_mi355x.h).

The classes mirror the reference's module interfaces:

* :class:`Grid` -- the set-up of ``initialize_model`` (grid points, halo
  tables, masks, metrics, multigrid hierarchy), computed on the host.
* :class:`Model` -- the device state and the time step ``step::step_dt`` +
  ``step::wrap_prognostics``; the reference operators, ``wrap_data::wrap``,
  ``elliptic_solver::elliptic`` and ``multigrid_solver::multigrid`` on host
  arrays in the reference layout.

Arrays are numpy float64 in Fortran order with the reference's shapes, e.g.
cell fields (iim, jjm, 10).  The library is built in-tree by
``build_lib.py``; importing this module raises if it is missing, there is
no fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libswm_mi355x.so")

SWM_FP64 = 0
SWM_MIXED = 1
# swm_set_fusion flags
FUSE_INJECT_SWEEP = 1
FUSE_HALO_PUSH = 2
FUSE_MASS_ANTI = 4
FUSE_ALL = 7
# advection of step_dt: l_fct, l_centered, l_upwind, l_remap
ADVECTION = {"fct": 0, "centered": 1, "upwind": 2, "remap": 3}

# physical constants of the reference build (utilities/layer_1/params_physical.F:45-52)
REF_RADIUS = 6.37122e6
REF_GRAV = 9.80616
REF_OMEGA = 7.292e-5

_STATUS = {0: "SWM_OK", 1: "SWM_ERR_ARG", 2: "SWM_ERR_DIMS", 3: "SWM_ERR_HIP", 4: "SWM_ERR_NOT_CONVERGED",
           5: "SWM_ERR_NAN", 6: "SWM_ERR_NO_DEVICE", 7: "SWM_ERR_STATE"}
globals().update({v: k for k, v in _STATUS.items()})


class SwmError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{_STATUS.get(status, status)}: {msg}")
        self.status = status


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libswm_mi355x.so not built ({LIB_PATH}); run shallow-water-model_amd/build_lib.py")
    return ctypes.CDLL(LIB_PATH)


lib = _load()

_c_int_p = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)


class _Config(ctypes.Structure):
    _fields_ = [("ntracer", ctypes.c_int), ("dt", ctypes.c_double), ("grav", ctypes.c_double),
                ("precision", ctypes.c_int), ("device", ctypes.c_int), ("advection", ctypes.c_int)]


class _Namelist(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("tau_end_day", "tau_end_hour", "dt")] + \
               [(n, ctypes.c_int) for n in ("l_restart", "l_upwind", "l_centered", "l_fct", "l_remap", "n_restart",
                                            "n_output", "n_diagnostics")]


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("swm_last_error", ctypes.c_char_p, [])
_sig("swm_grid_points", ctypes.c_int, [ctypes.c_int, ctypes.c_void_p])
_sig("swm_grid_create", ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                       ctypes.POINTER(ctypes.c_void_p)])
_sig("swm_grid_destroy", None, [ctypes.c_void_p])
_sig("swm_grid_dims", ctypes.c_int, [ctypes.c_void_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p])
_sig("swm_grid_array_len", ctypes.c_int64, [ctypes.c_void_p, ctypes.c_char_p])
_sig("swm_grid_array", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p])
_sig("swm_create", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_Config), ctypes.POINTER(ctypes.c_void_p)])
_sig("swm_destroy", None, [ctypes.c_void_p])
_sig("swm_init_state", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 4)
_sig("swm_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int])
_sig("swm_get_field", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p])
_sig("swm_synchronize", ctypes.c_int, [ctypes.c_void_p])
_sig("swm_step_count", ctypes.c_int, [ctypes.c_void_p])
_sig("swm_last_vcycles", ctypes.c_int, [ctypes.c_void_p])
_sig("swm_total_mass", ctypes.c_int, [ctypes.c_void_p, _dp])
_sig("swm_set_solver", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int])
_sig("swm_last_residual", ctypes.c_int, [ctypes.c_void_p, _dp, _dp])
_sig("swm_set_mixed_start", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int])
_sig("swm_budgets", ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int])
_sig("swm_solver_stats", ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_longlong)] * 3)
_sig("swm_profile", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int])
_sig("swm_timing", ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _c_int_p])
_sig("swm_init_synthetic", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double])
_sig("swm_time_seconds", ctypes.c_longlong, [ctypes.c_void_p])
_sig("swm_write_restart", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p])
_sig("swm_read_restart", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p])
_sig("swm_restart_name", ctypes.c_int, [ctypes.c_longlong, ctypes.c_char_p, ctypes.c_int])
_sig("swm_restart_file_write", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_longlong]
     + [ctypes.c_void_p] * 3)
_sig("swm_restart_file_read", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_longlong)] + [ctypes.c_void_p] * 3)
_sig("swm_swtc5_state", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double] + [ctypes.c_void_p] * 4)
_sig("swm_read_namelist", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_Namelist)])
_sig("swm_init_vorticity", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 3)
_sig("swm_adj_forward", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int])
_sig("swm_tl_run", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 8)
_sig("swm_adj_run", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 8)
_sig("swm_global_real_sum", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, _dp])
_sig("swm_global_real_sum_device", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 2 + [ctypes.c_int64,
                                                                                               ctypes.c_void_p])
_sig("swm_global_real_sum_noalloc", ctypes.c_int, [ctypes.c_void_p] * 2 + [ctypes.c_int64, ctypes.c_void_p,
                                                                              ctypes.c_int, ctypes.c_void_p])
_sig("swm_op_set_mass_fct", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 14 + [ctypes.c_void_p] * 2 +
     [ctypes.c_double] + [ctypes.c_void_p] * 12)
_sig("swm_op_set_tracer_fct", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 5 + [ctypes.c_double] +
     [ctypes.c_void_p] * 12)
_sig("swm_self_test", ctypes.c_int, [ctypes.c_int])
_sig("swm_restart_clock", ctypes.c_int, [ctypes.c_void_p])
_sig("swm_set_halo_shadow", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int])
_sig("swm_halo_shadow_stats", ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_longlong)] * 2)
_sig("swm_set_linearisation", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int])
_sig("swm_operator_bench", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp])
_sig("swm_adj_limiter_stats", ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_longlong)] * 2)
_sig("swm_adj_tracer_limiter_stats", ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_longlong)] * 2)
_sig("swm_init_vorticity_tracer", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 4)
_sig("swm_tl_run_tracer", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 10)
_sig("swm_adj_run_tracer", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 10)
_sig("swm_wrap", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p])
_sig("swm_elliptic", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p])
_sig("swm_multigrid", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p])
_sig("swm_create_members", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_Config), ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_void_p)])
_sig("swm_members", ctypes.c_int, [ctypes.c_void_p])
_sig("swm_ctx_layout", ctypes.c_int, [ctypes.c_void_p, _c_int_p, _c_int_p, _c_int_p])
_sig("swm_set_remap_limiter", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int])
_sig("swm_init_diss_op", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p])
_sig("swm_op_dissipation", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                          ctypes.c_void_p])
_sig("swm_op_remap_mass_and_tracers", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 5 +
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p])
_sig("swm_qp_name", ctypes.c_int, [ctypes.c_longlong, ctypes.c_int, ctypes.c_char_p, ctypes.c_int])
_sig("swm_qp_create", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_longlong, ctypes.c_char_p,
                                     ctypes.POINTER(ctypes.c_void_p)])
_sig("swm_qp_create_coords", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_longlong, ctypes.c_char_p]
     + [ctypes.c_void_p] * 4 + [ctypes.POINTER(ctypes.c_void_p)])
_sig("swm_qp_pack", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p])
_sig("swm_qp_pack_field", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p])
_sig("swm_qp_write", ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong])
_sig("swm_qp_records", ctypes.c_int, [ctypes.c_void_p])
_sig("swm_qp_path", ctypes.c_char_p, [ctypes.c_void_p])
_sig("swm_qp_destroy", None, [ctypes.c_void_p])
_sig("swm_set_mixed_smoother", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int])
_sig("swm_set_overlap", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int])
_sig("swm_set_fusion", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int])
_sig("swm_set_mixed_coarse", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int])
_sig("swm_select_member", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int])
_sig("swm_set_fixed_cycles", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int])
_sig("swm_solver_track", ctypes.c_int, [ctypes.c_void_p, _dp])
_sig("swm_init_synthetic_member", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                 ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                                 ctypes.c_int, ctypes.c_double])


class _Choice(ctypes.Structure):
    _fields_ = [("concurrent", ctypes.c_int), ("steps_per_graph", ctypes.c_int), ("blk_step", ctypes.c_int),
                ("blk_mg", ctypes.c_int), ("ms_per_step", ctypes.c_double)]


_sig("swm_ensemble_create", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), _c_int_p, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_void_p)])
_sig("swm_ensemble_destroy", None, [ctypes.c_void_p])
_sig("swm_ensemble_set", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_Choice)])
_sig("swm_ensemble_get", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_Choice)])
_sig("swm_ensemble_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int])
_sig("swm_ensemble_tune", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(_Choice), ctypes.c_int,
                                         _c_int_p])

# reference operators (operators.F): name -> (integer arguments after ctx,
# array arguments as (name, leading shape), output (name, leading shape));
# the C entry points take the subroutine's argument list plus nsdm
_FIXED = {"npad": 1, "ndata": 3, "nspace": 2, "ncorners": 2, "nedges": 3, "nweights": 7, "r3": 3}
OPERATORS = {
    "curl_of_a_vector": (("npad", "nx", "ny", "nsdm", "ndata", "nspace", "ncorners"),
                         (("area_inv", ()), ("v_weights", (3, 2, 2)), ("xin", (2, 2))), ("xout", ())),
    "div_of_a_vector": (("npad", "nx", "ny", "nsdm", "ndata", "nspace", "ncorners"),
                        (("area_inv", ()), ("v_weights", (3, 2, 2)), ("xin", (2, 2))), ("xout", ())),
    "div_of_an_edge_flux": (("npad", "nx", "ny", "nsdm", "nedges"), (("area_inv", ()), ("xin", (3,))), ("xout", ())),
    "laplacian_of_a_scalar": (("npad", "nx", "ny", "nsdm", "nweights"), (("l_weights", (7,)), ("xin", ())),
                              ("xout", ())),
    "k_cross_v": (("nx", "ny", "nsdm", "nspace", "ncorners"), (("xin", (2, 2)),), ("xout", (2, 2))),
    "gradient": (("npad", "nx", "ny", "nsdm", "ndata", "nspace", "ncorners"),
                 (("corner_area_inv", (2,)), ("v_weights", (3, 2, 2)), ("xin", ())), ("xout", (2, 2))),
    "average_to_cell_corners": (("npad", "nx", "ny", "nsdm", "ndata", "ncorners"),
                                (("s_weights", (3, 2)), ("xin", ())), ("xout", (2,))),
    "calculate_center_kinetic_energy": (("npad", "nx", "ny", "nsdm", "ndata", "nspace", "ncorners"),
                                        (("sub_area", (3, 2)), ("area_inv", ()), ("ke_c", (2,))), ("ke", ())),
    "calculate_velocity": (("npad", "nx", "ny", "nsdm", "ndata", "nspace", "ncorners"),
                           (("corner_area_inv", (2,)), ("v_weights", (3, 2, 2)), ("psi", ()), ("chi", ())),
                           ("xout", (2, 2))),
    "centered_mass_flux": (("npad", "nx", "ny", "nsdm", "ndata", "ncorners", "nedges", "nspace", "r3"),
                           (("dc", (3, 2)), ("normal", (3, 3, 2)), ("mass_c", (2,)), ("u", (2, 2)), ("mass", ())),
                           ("flux", (3,))),
    "upstream_mass_flux": (("npad", "nx", "ny", "nsdm", "ndata", "ncorners", "nedges", "nspace", "r3"),
                           (("dc", (3, 2)), ("normal", (3, 3, 2)), ("u", (2, 2)), ("mass", ())), ("flux", (3,))),
    "centered_vorticity_flux": (("npad", "nx", "ny", "nsdm", "nedges"),
                                (("mass_flux", (3,)), ("mass", ()), ("xin", ())), ("xout", (3,))),
    "centered_divergence_curl": (("npad", "nx", "ny", "nsdm", "ndata", "ncorners", "nedges", "nspace", "r3"),
                                 (("dc", (3, 2)), ("tangent", (3, 3, 2)), ("mass_c", (2,)), ("eta", ()),
                                  ("u", (2, 2)), ("mass", ())), ("flux", (3,))),
}
for _n, (_ints, _ins, _o) in OPERATORS.items():
    _extra = [ctypes.c_int, ctypes.c_int] if _n == "average_to_cell_corners" else []
    _sig("swm_op_" + _n, ctypes.c_int,
         [ctypes.c_void_p] + [ctypes.c_int] * len(_ints) + [ctypes.c_void_p] * (len(_ins) + 1) + _extra)


def _check(status):
    if status != 0:
        raise SwmError(status, lib.swm_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a, shape=None):
    a = np.asfortranarray(np.asarray(a, dtype=np.float64))
    if shape is not None and a.shape != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {a.shape}")
    return a


def grid_points(g):
    """Cell centres (3, 10*4^g+2) of the level-g icosahedral grid, global order."""
    n = 10 * 4 ** g + 2
    out = np.zeros((3, n), dtype=np.float64, order="F")
    _check(lib.swm_grid_points(g, _ptr(out)))
    return out


def restart_name(nsecond):
    """write_swm_restart's file name for a model time (swm_restart.F:122-145)."""
    buf = ctypes.create_string_buffer(64)
    _check(lib.swm_restart_name(int(nsecond), buf, 64))
    return buf.value.decode()


def read_namelist(path):
    """The reference driver's data/namelist (model_control.F:98-139) as a dict."""
    nl = _Namelist()
    _check(lib.swm_read_namelist(os.fsencode(path), ctypes.byref(nl)))
    return {n: getattr(nl, n) for n, _ in _Namelist._fields_}


def swtc5_state(grid, ntracer=5, grav=REF_GRAV):
    """Shallow-water test case 5 (initialize.F:330-513): dict(h, hs, u, tracer)
    in the reference layout, halos filled."""
    s = (grid.iim, grid.jjm, 10)
    out = {"h": np.zeros(s, order="F"), "hs": np.zeros(s, order="F"), "u": np.zeros((2, 2) + s, order="F"),
           "tracer": np.zeros((grid.iim, grid.jjm, ntracer, 10), order="F")}
    _check(lib.swm_swtc5_state(grid.handle, ntracer, grav, _ptr(out["h"]), _ptr(out["hs"]), _ptr(out["u"]),
                               _ptr(out["tracer"])))
    return out


def write_restart_file(grid, path, h, u, tracer, nsecond=0):
    """Writes a restart file (io_utilities.F direct-access format, as
    swm_restart.F:113-175) from host arrays in the reference layout."""
    s = (grid.iim, grid.jjm, 10)
    tracer = np.asarray(tracer)
    nt = tracer.shape[2]
    h, u, tr = _f64(h, s), _f64(u, (2, 2) + s), _f64(tracer, (grid.iim, grid.jjm, nt, 10))
    _check(lib.swm_restart_file_write(grid.handle, os.fsencode(path), nt, int(nsecond), _ptr(h), _ptr(u), _ptr(tr)))


def read_restart_file(grid, path, ntracer):
    """Reads a restart file into (nsecond, h, u, tracer), every slot filled from
    its global cell's record (swm_restart.F:25-45)."""
    s = (grid.iim, grid.jjm, 10)
    h = np.zeros(s, order="F")
    u = np.zeros((2, 2) + s, order="F")
    tr = np.zeros((grid.iim, grid.jjm, ntracer, 10), order="F")
    ts = ctypes.c_longlong()
    _check(lib.swm_restart_file_read(grid.handle, os.fsencode(path), ntracer, ctypes.byref(ts), _ptr(h), _ptr(u),
                                     _ptr(tr)))
    return ts.value, h, u, tr


QP_FIELDS = ("thickness", "height", "relative", "divergence", "absolute", "kinetic_energy",
             "tracer_1", "tracer_2", "tracer_3", "tracer_4", "tracer_5")


def qp_name(nsecond=0, task=0):
    """initialize_swm_qp's file name (swm_qp_module.F:130-136, :176)."""
    buf = ctypes.create_string_buffer(64)
    _check(lib.swm_qp_name(int(nsecond), int(task), buf, 64))
    return buf.value.decode()


class QpFile:
    """The quick-plot netCDF file of swm_qp_module.F: initialize_swm_qp at
    construction, pack_swm_qp (pack / pack_field) and write_swm_qp (write).

    path None uses the reference's name under qp_output/ (the directory must
    exist); coords = (center_lat, center_lon, corner_lat, corner_lon) in the
    reference layout replaces the coordinates derived from the grid."""

    def __init__(self, grid, simulation_tag="SWTC#5", nsecond=0, path=None, coords=None):
        h = ctypes.c_void_p()
        p = None if path is None else os.fsencode(path)
        tag = simulation_tag.encode()
        if coords is None:
            _check(lib.swm_qp_create(grid.handle, tag, int(nsecond), p, ctypes.byref(h)))
        else:
            s = (grid.iim, grid.jjm, 10)
            clat, clon = _f64(coords[0], s), _f64(coords[1], s)
            crlat, crlon = _f64(coords[2], (2,) + s), _f64(coords[3], (2,) + s)
            _check(lib.swm_qp_create_coords(grid.handle, tag, int(nsecond), p, _ptr(clat), _ptr(clon), _ptr(crlat),
                                            _ptr(crlon), ctypes.byref(h)))
        self._h = h
        self.grid = grid

    @property
    def path(self):
        return lib.swm_qp_path(self._h).decode()

    @property
    def records(self):
        return lib.swm_qp_records(self._h)

    def pack(self, model):
        """pack_swm_qp: the model's current time level."""
        _check(lib.swm_qp_pack(self._h, model._h))

    def pack_field(self, name, x):
        """qpsum (replace) of one field from a host array (iim, jjm, 10)."""
        x = _f64(x, (self.grid.iim, self.grid.jjm, 10))
        _check(lib.swm_qp_pack_field(self._h, name.encode(), _ptr(x)))

    def write(self, nsecond):
        """write_swm_qp: one record at model time nsecond; the packed fields
        return to zero."""
        _check(lib.swm_qp_write(self._h, int(nsecond)))

    def close(self):
        if self._h:
            lib.swm_qp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# shapes of the named grid arrays, in units of (iim, jjm, 10) slots
_GRID_SHAPES = {
    "p": (3,), "c": (3, 2), "w": (3, 3, 2), "trans_to_X1": (3, 3, 2), "trans_to_X2": (3, 3, 2),
    "area": (), "area_inv": (), "corner_area": (2,), "corner_area_inv": (2,),
    "sub_area": (3, 2), "dc": (3, 2), "normal": (3, 3, 2), "tangent": (3, 3, 2), "v_weights": (3, 2, 2),
    "s_weights": (3, 2), "l_weights": (7,), "f": (), "grid_center_mask": (), "grid_corner_mask": (2,),
    "polygon_type": (), "grid_center_index": (), "grid_corner_index": (2,),
    "remap_vertex_mask": (6,), "remap_num_vertices": (), "remap_trans_X1_to_X3": (3, 3),
    "remap_trans_X1_to_X4": (3, 3, 3), "remap_trans_X4_to_X3": (2, 2, 4, 3), "remap_cl_x4": (2, 3),
    "remap_cr_x4": (2, 3), "remap_tl_x4": (2, 3), "remap_tr_x4": (2, 3), "remap_bl_x4": (2, 3), "remap_br_x4": (2, 3),
    "remap_p_x4": (2, 4, 3), "remap_c_x3": (2, 6), "remap_area_x3": (), "remap_area_inv_x3": (), "remap_xav": (),
    "remap_yav": (), "remap_xxav": (), "remap_xyav": (), "remap_yyav": (), "remap_xxxav": (), "remap_xxyav": (),
    "remap_xyyav": (), "remap_yyyav": (), "remap_gwgt_x3": (3, 7), "remap_garea_inv": (),
}
_INT_ARRAYS = {"grid_center_index", "grid_corner_index", "face_dst", "face_src", "vrtx_dst", "vrtx_src",
               "edge_dst", "edge_src"}


class Grid:
    """Host-side grid set-up at icosahedral level g (10 panels of 2^g x 2^g cells
    plus the two poles, one-cell halos)."""

    def __init__(self, g, p_list=None, radius=REF_RADIUS, omega=REF_OMEGA):
        h = ctypes.c_void_p()
        pl = None if p_list is None else _f64(p_list, (3, 10 * 4 ** g + 2))
        _check(lib.swm_grid_create(g, _ptr(pl), radius, omega, ctypes.byref(h)))
        self._h = h
        gi, iim, jjm, mig = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib.swm_grid_dims(h, ctypes.byref(gi), ctypes.byref(iim), ctypes.byref(jjm), ctypes.byref(mig)))
        self.g, self.iim, self.jjm, self.max_ig = gi.value, iim.value, jjm.value, mig.value
        self.radius, self.omega = radius, omega

    @property
    def handle(self):
        return self._h

    def array(self, name):
        n = lib.swm_grid_array_len(self._h, name.encode())
        if n < 0:
            raise KeyError(name)
        dt = np.int32 if name in _INT_ARRAYS else np.float64
        out = np.zeros(n, dtype=dt)
        _check(lib.swm_grid_array(self._h, name.encode(), _ptr(out)))
        if name in _GRID_SHAPES:
            return out.reshape(_GRID_SHAPES[name] + (self.iim, self.jjm, 10), order="F")
        if name.startswith("lap_w"):
            m = 2 + 2 ** int(name[5:])
            return out.reshape((7, m, m, 10), order="F")
        if name == "tot_area":
            return float(out[0])
        return out

    def close(self):
        if self._h:
            lib.swm_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Groups of members (Models made with members=n, mixed precision with a
    fixed schedule) advanced together in captured HIP graphs; group k takes
    substeps[k] of its own steps per ensemble step."""

    def __init__(self, groups, substeps):
        self.groups = list(groups)
        arr = (ctypes.c_void_p * len(self.groups))(*[g._h.value for g in self.groups])
        sub = (ctypes.c_int * len(self.groups))(*substeps)
        h = ctypes.c_void_p()
        _check(lib.swm_ensemble_create(arr, sub, len(self.groups), ctypes.byref(h)))
        self._h = h

    def step(self, nsteps=1):
        _check(lib.swm_ensemble_step(self._h, nsteps))

    def choice(self):
        """Launch choice and the last measured ms per ensemble step."""
        c = _Choice()
        _check(lib.swm_ensemble_get(self._h, ctypes.byref(c)))
        return {k: getattr(c, k) for k, _ in _Choice._fields_}

    def set_choice(self, concurrent=0, steps_per_graph=1, blk_step=256, blk_mg=256):
        c = _Choice(concurrent, steps_per_graph, blk_step, blk_mg, 0.0)
        _check(lib.swm_ensemble_set(self._h, ctypes.byref(c)))

    def tune(self, rounds=2, max_trials=32):
        """Autotune the launch choice; returns the trials (dicts with
        ms_per_step)."""
        tr = (_Choice * max_trials)()
        n = ctypes.c_int()
        _check(lib.swm_ensemble_tune(self._h, rounds, tr, max_trials, ctypes.byref(n)))
        return [{k: getattr(t, k) for k, _ in _Choice._fields_} for t in tr[:min(n.value, max_trials)]]

    def close(self):
        if self._h:
            lib.swm_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Model:
    """Device-resident model state on one GPU (the reference's prognostics
    module plus the saved state of step.F and the elliptic warm start)."""

    def __init__(self, grid, ntracer=1, dt=450.0, grav=REF_GRAV, precision=SWM_FP64, device=0, advection="fct",
                 members=1):
        cfg = _Config(ntracer, dt, grav, precision, device, ADVECTION[advection])
        h = ctypes.c_void_p()
        if members == 1:
            _check(lib.swm_create(grid.handle, ctypes.byref(cfg), ctypes.byref(h)))
        else:
            _check(lib.swm_create_members(grid.handle, ctypes.byref(cfg), members, ctypes.byref(h)))
        self._h = h
        self.grid, self.ntracer, self.dt = grid, ntracer, dt
        self.iim, self.jjm = grid.iim, grid.jjm
        self.members = members

    def select(self, member):
        """Member addressed by the host-side calls (init, field, budgets,
        restart files, operators) from now on."""
        _check(lib.swm_select_member(self._h, member))
        return self

    def set_overlap(self, enable=True):
        """Tracer transport on a second stream beside the dynamics (FCT):
        False/0 off, True/1 from the mass update on, n >= 2 started by the
        first V-cycle at level g - n + 1."""
        _check(lib.swm_set_overlap(self._h, int(enable)))

    def set_mixed_coarse(self, direct=True, level=3):
        """Mixed path: the coarsest levels as one dense solve of level 3 or
        of level 4 (where level 5 exists) (direct=True, level 3 or 4), or the
        reference's V-cycle below level 4 (direct=False)."""
        if level not in (3, 4):
            raise ValueError("level is 3 or 4")
        _check(lib.swm_set_mixed_coarse(self._h, (2 if level == 4 else 1) if direct else 0))

    def set_fusion(self, flags=None):
        """Kernel fusions of the multigrid cycle (FUSE_* flags; None: all).
        The results are the same bit for bit either way."""
        _check(lib.swm_set_fusion(self._h, FUSE_ALL if flags is None else int(flags)))

    def set_mixed_smoother(self, fine_sweeps=2, half_weights=True):
        """Mixed path: finest-level sweeps before / after the coarse
        correction (1 to 4), and half-precision weights in the smoother and
        restriction above the three coarsest levels."""
        _check(lib.swm_set_mixed_smoother(self._h, fine_sweeps, 1 if half_weights else 0))

    def set_fixed_cycles(self, ncycles, nstart=8):
        """Mixed solve: exactly ncycles correction cycles per solve (the first
        two after an initialisation: max(ncycles, nstart)), no host test (0:
        the convergence test)."""
        _check(lib.swm_set_fixed_cycles(self._h, ncycles, nstart))

    def solver_track(self):
        """(max mean r^2 of psi, of chi, solves above the criterion) of the
        selected member under the fixed schedule."""
        out = (ctypes.c_double * 3)()
        _check(lib.swm_solver_track(self._h, out))
        return out[0], out[1], int(out[2])

    def init_synthetic_member(self, member_seed, vort_rel=0.01, seed_index=0, lmax=20, umax=40.0, h0=5000.0,
                              h_rms=50.0, nbumps=4, bump_radius=500.0e3):
        """Seeded synthetic state with a member perturbation of the vorticity
        (swm_init_synthetic_member), for the selected member."""
        _check(lib.swm_init_synthetic_member(self._h, seed_index, lmax, umax, h0, h_rms, nbumps, bump_radius,
                                             member_seed, vort_rel))

    def init_state(self, h, u, tracer, hs=None):
        s = (self.iim, self.jjm, 10)
        h = _f64(h, s)
        u = _f64(u, (2, 2) + s)
        tr = _f64(tracer, (self.iim, self.jjm, self.ntracer, 10))
        hs = None if hs is None else _f64(hs, s)
        _check(lib.swm_init_state(self._h, _ptr(h), _ptr(hs), _ptr(u), _ptr(tr)))

    def step(self, nsteps=1):
        _check(lib.swm_step(self._h, nsteps))

    def field(self, name):
        s = (self.iim, self.jjm, 10)
        shape = {"u": (2, 2) + s, "tracer": (self.iim, self.jjm, self.ntracer, 10)}.get(name, s)
        out = np.zeros(shape, dtype=np.float64, order="F")
        _check(lib.swm_get_field(self._h, name.encode(), _ptr(out)))
        return out

    def synchronize(self):
        _check(lib.swm_synchronize(self._h))

    @property
    def step_count(self):
        return lib.swm_step_count(self._h)

    @property
    def last_vcycles(self):
        return lib.swm_last_vcycles(self._h)

    def op(self, name, *arrays, flag=True, min_value=None):
        """One reference operator (``operators::<name>``) on host arrays in the
        reference layout, metric arrays included, in the subroutine's
        argument order: (leading dims, nx, ny) for one panel or (leading
        dims, nx, ny, nsdm) for nsdm panels at once."""
        ints, ins, (_, olead) = OPERATORS[name]
        if len(arrays) != len(ins):
            raise TypeError(f"{name} takes {len(ins)} arrays: {[n for n, _ in ins]}")
        lead0 = ins[0][1]
        a0 = np.asarray(arrays[0])
        nx, ny = a0.shape[len(lead0)], a0.shape[len(lead0) + 1]
        nsdm = a0.shape[len(lead0) + 2] if a0.ndim == len(lead0) + 3 else 1
        tail = (nx, ny) if a0.ndim == len(lead0) + 2 else (nx, ny, nsdm)
        arrs = [_f64(a, lead + tail) for a, (_, lead) in zip(arrays, ins)]
        vals = dict(_FIXED, nx=nx, ny=ny, nsdm=nsdm)
        out = np.zeros(olead + tail, dtype=np.float64, order="F")
        extra = []
        if name == "average_to_cell_corners":
            extra = [1 if flag else 0, -1 if min_value is None else (1 if min_value else 0)]
        _check(getattr(lib, "swm_op_" + name)(self._h, *[vals[k] for k in ints], *[_ptr(a) for a in arrs],
                                              _ptr(out), *extra))
        return out

    def set_mass_fct(self, n3, n4, tend_index, tend_weights, area_inv, area, dc, normal, mss, mss_c, u, msh_f):
        """mass_fct::set_mass_fct (mass_fct.F:23-539) on host arrays in the
        reference layout, whole sphere: mss (iim,jjm,1,2,10) and msh_f
        (3,iim,jjm,1,3,10) are updated in place; returns (mss_low, mss_flux_horz)."""
        s = (self.iim, self.jjm)
        for a, shp in ((mss, s + (1, 2, 10)), (msh_f, (3,) + s + (1, 3, 10))):
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.f_contiguous and a.shape == shp):
                raise TypeError(f"expected a Fortran-ordered float64 array of shape {shp}")
        ins = [_f64(area_inv, s + (10,)), _f64(area, s + (10,)), _f64(dc, (3, 2) + s + (10,)),
               _f64(normal, (3, 3, 2) + s + (10,))]
        mc, uu = _f64(mss_c, (2,) + s + (1, 10)), _f64(u, (2, 2) + s + (1, 10))
        ti = np.ascontiguousarray(tend_index, dtype=np.int32)
        tw = np.ascontiguousarray(tend_weights, dtype=np.float64)
        low = np.zeros(s + (1, 10), order="F")
        flux = np.zeros((3,) + s + (1, 10), order="F")
        _check(lib.swm_op_set_mass_fct(self._h, self.iim, self.jjm, 1, 10, 2, 3, 2, 3, 3, 2, 1, 3, n3, n4, _ptr(ti),
                                       _ptr(tw), self.dt, *[_ptr(a) for a in ins], _ptr(mss), _ptr(mc), _ptr(uu),
                                       _ptr(low), _ptr(flux), _ptr(msh_f), None, None))
        return low, flux

    def set_tracer_fct(self, area_inv, area, trc, trc_high, mss, mss_low, mss_horz, mss_flux_horz):
        """tracer_fct::set_tracer_fct (tracer_fct.F:22-670) on host arrays,
        whole sphere; returns (trc_fct, trc_flux_horz, msstrc_low_order)."""
        s = (self.iim, self.jjm)
        ins = [_f64(a, s + (1, 10)) for a in (trc, trc_high, mss, mss_low, mss_horz)]
        fl = _f64(mss_flux_horz, (3,) + s + (1, 10))
        out, tf, ml = np.zeros(s + (1, 10), order="F"), np.zeros((3,) + s + (1, 10), order="F"), \
            np.zeros(s + (1, 10), order="F")
        _check(lib.swm_op_set_tracer_fct(self._h, self.iim, self.jjm, 1, 10, 3, self.dt, _ptr(_f64(area_inv, s + (10,))),
                                         _ptr(_f64(area, s + (10,))), *[_ptr(a) for a in ins], _ptr(fl), _ptr(out),
                                         _ptr(tf), _ptr(ml), None))
        return out, tf, ml

    def wrap(self, kind, x, nspace=1):
        """Halo fill (wrap_data::wrap) of x in place; kind "face" x is
        (iim,jjm[,nlayer],10), "vrtx" (nspace,2,iim,jjm[,nlayer],10), "edge"
        (3,iim,jjm[,nlayer],10).  Returns x."""
        k = {"face": 0, "vrtx": 1, "edge": 2}[kind]
        lead = {0: 1, 1: 2 * nspace, 2: 3}[k]
        if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.f_contiguous):
            raise TypeError("wrap works in place on a Fortran-ordered float64 array")
        nl = x.size // (lead * self.iim * self.jjm * 10)
        if nl * lead * self.iim * self.jjm * 10 != x.size:
            raise ValueError("array size does not match the grid")
        _check(lib.swm_wrap(self._h, k, nspace, nl, _ptr(x)))
        return x

    def elliptic(self, rhs, x0):
        """elliptic_solver::elliptic on (iim,jjm,nvar,10) arrays; returns
        (rhs less its area-weighted mean, solution)."""
        r = np.array(rhs, dtype=np.float64, order="F", copy=True)
        x = np.array(x0, dtype=np.float64, order="F", copy=True)
        if r.shape != x.shape or r.shape[:2] != (self.iim, self.jjm) or r.shape[3] != 10:
            raise ValueError("expected (iim, jjm, nvar, 10) arrays")
        _check(lib.swm_elliptic(self._h, r.shape[2], _ptr(r), _ptr(x)))
        return r, x

    def multigrid(self, beta, alpha):
        """One V-cycle (multigrid_solver::multigrid) on (iim,jjm,nvar,10)
        arrays; returns the new alpha."""
        b = _f64(beta)
        a = np.array(alpha, dtype=np.float64, order="F", copy=True)
        if b.shape != a.shape or b.shape[:2] != (self.iim, self.jjm) or b.shape[3] != 10:
            raise ValueError("expected (iim, jjm, nvar, 10) arrays")
        _check(lib.swm_multigrid(self._h, b.shape[2], _ptr(b), _ptr(a)))
        return a

    def set_solver(self, tol=1.0e-20, relative=False, max_cycles=12):
        """Convergence settings of the elliptic solve; the defaults are the
        reference's (mean squared residual <= 1e-20 within 12 V-cycles)."""
        _check(lib.swm_set_solver(self._h, tol, 1 if relative else 0, max_cycles))

    def set_mixed_start(self, extrapolate=2, min_cycles=1):
        """Mixed path: the warm start, extrapolated from the last three
        solutions (2), the last two (1) or the last one (0, False), and the
        correction cycles before the first convergence test."""
        _check(lib.swm_set_mixed_start(self._h, int(extrapolate), min_cycles))

    def last_residual(self):
        """(mean squared residuals, mean squared right-hand sides) of psi and
        chi at the end of the last elliptic solve."""
        r = (ctypes.c_double * 2)()
        b = (ctypes.c_double * 2)()
        _check(lib.swm_last_residual(self._h, r, b))
        return list(r), list(b)

    def budgets(self):
        """compute_budgets global averages (budgets.F:31-247): mass, ke, pe,
        tf, te, potential enstrophy, eta, div, then one mass-weighted variance
        per tracer (at most 16)."""
        n = 8 + min(self.ntracer, 16)
        out = (ctypes.c_double * n)()
        _check(lib.swm_budgets(self._h, out, n))
        return np.array(out)

    def solver_stats(self):
        """(elliptic solves, V-cycles, residual checks) since initialisation."""
        v = [ctypes.c_longlong() for _ in range(3)]
        _check(lib.swm_solver_stats(self._h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def init_synthetic(self, seed_index=0, lmax=20, umax=40.0, h0=5000.0, h_rms=50.0, nbumps=4,
                       bump_radius=500.0e3):
        """Seeded synthetic initial state built on the device (swm_init_synthetic)."""
        _check(lib.swm_init_synthetic(self._h, seed_index, lmax, umax, h0, h_rms, nbumps, bump_radius))

    def profile(self, enable=True, every=1):
        """HIP-event timing of the finest-level smoother: every `every`-th
        launch while enabled."""
        _check(lib.swm_profile(self._h, max(1, int(every)) if enable else 0))

    def timing(self):
        """(device ms of the last step call, mean ms, bytes and launch count of
        the profiled kernel: the finest-level smoother)."""
        a, b, c, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        _check(lib.swm_timing(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(n)))
        return a.value, b.value, c.value, n.value

    # ---------------------------------------------------------- TL / AD
    def init_vorticity(self, zeta, delta, h, tracer=None):
        """Initial state from relative vorticity, divergence and thickness
        (iim,jjm,10): streamfunction and velocity potential solved, u from
        them, the rest derived as by init_state; tracers (iim,jjm,ntracer,10)
        as given, or zero."""
        s = (self.iim, self.jjm, 10)
        z, d, hh = _f64(zeta, s), _f64(delta, s), _f64(h, s)
        if tracer is None:
            _check(lib.swm_init_vorticity(self._h, _ptr(z), _ptr(d), _ptr(hh)))
        else:
            t = _f64(tracer, (self.iim, self.jjm, self.ntracer, 10))
            _check(lib.swm_init_vorticity_tracer(self._h, _ptr(z), _ptr(d), _ptr(hh), _ptr(t)))

    def adj_forward(self, nsteps, every=16):
        """The window of the linearised model: nsteps steps from the freshly
        initialised state, snapshots every `every` steps."""
        _check(lib.swm_adj_forward(self._h, nsteps, every))

    def _state_arrays(self):
        s = (self.iim, self.jjm, 10)
        return [np.zeros(s, order="F"), np.zeros(s, order="F"), np.zeros(s, order="F"),
                np.zeros((2, 2) + s, order="F")]

    def tl_run(self, dh=None, deta=None, ddiv=None, du=None, control=None, init_only=False, dtracer=None):
        """Tangent-linear model of the window: the perturbation of the state
        read by the first step (dh, deta, ddiv, du) or of the control
        (control=(dzeta, ddelta, dh)); returns (h, eta, div, u) after the
        last step.  With dtracer (iim,jjm,ntracer,10), the perturbation of
        the tracers read by the first step (or of the initial tracers), the
        tracers are carried too and their perturbation is returned last."""
        s = (self.iim, self.jjm, 10)
        out = self._state_arrays()
        if control is not None:
            ins = [_f64(x, s) for x in control] + [np.zeros((2, 2) + s, order="F")]
            mode = 2 if init_only else 1
        else:
            ins = [_f64(dh, s), _f64(deta, s), _f64(ddiv, s), _f64(du, (2, 2) + s)]
            mode = 0
        if dtracer is None:
            _check(lib.swm_tl_run(self._h, mode, *[_ptr(x) for x in ins], *[_ptr(x) for x in out]))
            return tuple(out)
        ts = (self.iim, self.jjm, self.ntracer, 10)
        tin = _f64(dtracer, ts)
        tout = np.zeros(ts, order="F")
        _check(lib.swm_tl_run_tracer(self._h, mode, *[_ptr(x) for x in ins], _ptr(tin), *[_ptr(x) for x in out],
                                     _ptr(tout)))
        return tuple(out) + (tout,)

    def adj_run(self, lh, leta=None, ldiv=None, lu=None, control=False, init_only=False, ltracer=None):
        """Adjoint of the window for the adjoint (gradient) of the state
        after the last step; returns the adjoint of the state read by the
        first step (h, eta, div, u), or with control=True the gradient with
        respect to (zeta, delta, h).  With ltracer (iim,jjm,ntracer,10), the
        adjoint of the tracers after the last step, the tracers are carried
        too and their adjoint (the gradient with respect to the initial
        tracers) is returned last."""
        s = (self.iim, self.jjm, 10)
        z = np.zeros(s, order="F")
        ins = [_f64(lh, s), _f64(z if leta is None else leta, s), _f64(z if ldiv is None else ldiv, s),
               _f64(np.zeros((2, 2) + s) if lu is None else lu, (2, 2) + s)]
        out = self._state_arrays()
        mode = 2 if init_only else (1 if control else 0)
        if ltracer is None:
            _check(lib.swm_adj_run(self._h, mode, *[_ptr(x) for x in ins], *[_ptr(x) for x in out]))
            extra = ()
        else:
            ts = (self.iim, self.jjm, self.ntracer, 10)
            tin = _f64(ltracer, ts)
            tout = np.zeros(ts, order="F")
            _check(lib.swm_adj_run_tracer(self._h, mode, *[_ptr(x) for x in ins], _ptr(tin),
                                          *[_ptr(x) for x in out], _ptr(tout)))
            extra = (tout,)
        if init_only or control:
            return tuple(out[:3]) + extra
        return tuple(out) + extra

    def global_real_sum(self, x, mask=None):
        """global_real_sum (global_reductions.F:77-98): sum(x * mask) of host
        arrays, on the device with a compensated final sum."""
        xf = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1, order="F"))
        mf = None if mask is None else np.ascontiguousarray(np.asarray(mask, dtype=np.float64).reshape(-1, order="F"))
        if mf is not None and mf.size != xf.size:
            raise ValueError("mask and x differ in size")
        out = ctypes.c_double()
        _check(lib.swm_global_real_sum(self._h, _ptr(xf), _ptr(mf), xf.size, ctypes.byref(out)))
        return out.value

    def set_halo_shadow(self, enable=True):
        """Check every halo fill against the host's fill with the same tables."""
        _check(lib.swm_set_halo_shadow(self._h, 1 if enable else 0))

    def halo_shadow_stats(self):
        """(fields checked, values that differed) of the halo shadow check."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib.swm_halo_shadow_stats(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def restart_clock(self):
        """Restart the step counter of every member, keeping its latest state:
        the next step is a forward Euler step (swm_restart_clock)."""
        _check(lib.swm_restart_clock(self._h))

    def set_linearisation(self, exact=True):
        """TL/AD of the window's solves: the forward's own V-cycles (True, the
        default: the derivative of the model as it runs) or converged solves
        (False)."""
        _check(lib.swm_set_linearisation(self._h, 1 if exact else 0))

    OPERATOR_BENCH = ("div_of_an_edge_flux", "curl_of_a_vector", "laplacian_of_a_scalar", "face_halo_fill",
                      "remap_transport", "remap_fields", "dissipation")

    def operator_bench(self, op, reps=200):
        """Microbench of one operator on the resident state (FP64): op is a
        name of OPERATOR_BENCH; returns (us per launch, algorithmic bytes
        per launch)."""
        us, b = ctypes.c_double(), ctypes.c_double()
        _check(lib.swm_operator_bench(self._h, self.OPERATOR_BENCH.index(op), reps, ctypes.byref(us),
                                      ctypes.byref(b)))
        return us.value, b.value

    def limiter_stats(self):
        """FCT window: (edge-steps with a frozen limiter coefficient below 1,
        edge-steps in all) over the base steps of the last TL or AD sweep."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib.swm_adj_limiter_stats(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def tracer_limiter_stats(self):
        """FCT window, tracer runs: (edge-steps with a frozen tracer limiter
        coefficient below 1, edge-steps in all, over the tracers) of the last
        TL or AD sweep with tracers."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib.swm_adj_tracer_limiter_stats(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    @property
    def time_seconds(self):
        """Model time in whole seconds (nsecond_dyn)."""
        return lib.swm_time_seconds(self._h)

    def set_remap_limiter(self, limit_min=True, limit_max=True):
        """Incremental remapping: the tracer limiter's l_limit_tracers_min /
        _max switches (swm_set_remap_limiter; both on by default)."""
        _check(lib.swm_set_remap_limiter(self._h, int(bool(limit_min)), int(bool(limit_max))))

    def init_diss_op(self):
        """init_diss_op (swm_dissipation_operator.F:50-203) for the model's
        grid."""
        _check(lib.swm_init_diss_op(self._h, self.grid.handle))

    def dissipation(self, norder, visc, xin):
        """dissipation (swm_dissipation_operator.F:212-443) of corner vectors
        (2, 2, iim, jjm, 10)."""
        x = _f64(xin, (2, 2, self.iim, self.jjm, 10))
        out = np.zeros_like(x, order="F")
        _check(lib.swm_op_dissipation(self._h, int(norder), float(visc), _ptr(x), _ptr(out)))
        return out

    def remap_mass_and_tracers(self, mass, trace, dt, uc):
        """remap_mass_and_tracers (remap.F:1072-1604) on host arrays, mass
        (iim, jjm, 10), trace (iim, jjm, ntrace, 10), uc (2, 2, iim, jjm, 10):
        returns (mass, trace, edgeflux (3, iim, jjm, 10))."""
        s = (self.iim, self.jjm, 10)
        m = np.array(_f64(mass, s), order="F")
        nt = np.asarray(trace).shape[2]
        t = np.array(_f64(trace, (self.iim, self.jjm, nt, 10)), order="F")
        u = _f64(uc, (2, 2) + s)
        ef = np.zeros((3,) + s, order="F")
        _check(lib.swm_op_remap_mass_and_tracers(self._h, self.iim, self.jjm, 1, 10, nt, _ptr(m), _ptr(t), float(dt),
                                                 _ptr(u), _ptr(ef)))
        return m, t, ef

    def write_restart(self, path=None):
        """write_swm_restart: the current state and model time; path None uses
        the reference's name under restarts/."""
        _check(lib.swm_write_restart(self._h, None if path is None else os.fsencode(path)))

    def read_restart(self, path=None):
        """read_swm_restart: state and model time from a file (hs unchanged);
        the next step is a forward step, as after any initialisation."""
        _check(lib.swm_read_restart(self._h, None if path is None else os.fsencode(path)))

    def total_mass(self):
        m = ctypes.c_double()
        _check(lib.swm_total_mass(self._h, ctypes.byref(m)))
        return m.value

    def close(self):
        if self._h:
            lib.swm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
