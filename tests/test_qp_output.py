"""Quick-plot netCDF output (swm/source/swm_qp_module.F over
utilities/layer_1/netcdf_geodesic.F and netcdf_general_tools.F), on the host.

The file is read back with scipy.io.netcdf_file, a netCDF classic reader
written independently of the library's writer: it locates every variable by
the offsets in the header, so a wrong header size, slab size, record size or
padding reads wrong values.  Checked against the reference's definitions:
  - dimensions, global and variable attributes, types and shapes as
    ginit_nc / define_time / define_geo_grid / nc_init_geo define them;
  - the fields: the FP64 values of the real cells (grid_center_mask > 0.5)
    in panel, j, i order, rounded to FP32; fields not packed since the last
    write are zero (write_swm_qp resets them);
  - the time variables of set_time: the period since the previous write in
    days, its middle as "time", char_time the date in the no-leap calendar,
    checked against Python's calendar for a non-leap year;
  - the coordinates by what they must be geometrically: interp_indx is a
    permutation of the global cell numbers and each centre is that global
    cell's generated point; each written corner is a corner of the grid
    whose three nearest centres include the cell, the corners run
    counter-clockwise, exactly the 12 pentagons repeat their fifth corner,
    and the cells' polygons tile the sphere (areas sum to 4 pi);
  - the reference's file name and the errors (unknown field, unwritable
    path).
Needs no GPU.
"""
import datetime
import os
import re

import numpy as np
import pytest
from scipy.io import netcdf_file
from scipy.spatial import cKDTree

import swm_mi355x as S


def _read(path):
    return netcdf_file(str(path), "r", mmap=False)


def _packed(grid, x):
    m = (grid.array("grid_center_mask") > 0.5).ravel(order="F")
    return np.asarray(x).ravel(order="F")[m]


def _xyz(lat, lon):
    return np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], -1)


@pytest.fixture(scope="module")
def grid3():
    return S.Grid(3)


def test_qp_file_definitions(grid3, tmp_path):
    g = grid3
    q = S.QpFile(g, "SWTC#5", 0, str(tmp_path / "qp.nc"))
    q.write(0)
    nc = _read(tmp_path / "qp.nc")
    n = 10 * 4 ** 3 + 2
    assert nc.dimensions == {"time": None, "char_len": 10, "grid_cells": n, "grid_size": n, "grid_corners": 6,
                             "grid_rank": 1}
    a = nc._attributes
    assert list(a) == ["calendar", "institution", "history", "run", "grid", "version", "extravaratts",
                       "total_grid_size"]
    assert (a["calendar"], a["institution"], a["run"], a["grid"], a["version"], a["extravaratts"]) == \
        (b"noleap", b"Colorado State University", b"SWTC#5", b"geodesic", b"0.6", b"yes")
    assert int(a["total_grid_size"]) == g.max_ig
    assert re.fullmatch(rb"Created: \d\d/\d\d/\d{4} at \d\d:\d\d [+-]\d{4} GMT ", a["history"])
    v = nc.variables
    names = ["time", "char_time", "start_period", "end_period", "period_length", "grid_center_lat",
             "grid_center_lon", "grid_dims", "grid_corner_lat", "grid_corner_lon"] + list(S.QP_FIELDS) + \
        ["interp_indx"]
    assert list(v) == names
    expect = {
        "time": (("time",), ">f8", {"quantity": b"time", "units": b"days since 1-1-1", "calendar": b"noleap"}),
        "char_time": (("time", "char_len"), "|S1", {"format": b"mm/dd/yyyy"}),
        "start_period": (("time",), ">f8", {"long_name": b"start of averaged period",
                                            "units": b"days since 1-1-1"}),
        "end_period": (("time",), ">f8", {"long_name": b"end of averaged period", "units": b"days since 1-1-1"}),
        "period_length": (("time",), ">f8", {"long_name": b"length of averaged period", "units": b"days"}),
        "grid_center_lat": (("grid_cells",), ">f8", {"units": b"radians"}),
        "grid_center_lon": (("grid_cells",), ">f8", {"units": b"radians"}),
        "grid_dims": (("grid_rank",), ">i4", {}),
        "grid_corner_lat": (("grid_cells", "grid_corners"), ">f8", {"units": b"radians"}),
        "grid_corner_lon": (("grid_cells", "grid_corners"), ">f8", {"units": b"radians"}),
        "interp_indx": (("grid_size",), ">i4", {}),
    }
    for k, (dims, dt, att) in expect.items():
        assert v[k].dimensions == dims and v[k].data.dtype.str == dt and dict(v[k]._attributes) == att, k
    # set_attribute_data (swm_qp_module.F:379-509)
    descr = {"thickness": (b"thickness of fliud layer", b"meters"),
             "height": (b"surface height of fliud layer", b"meters"),
             "relative": (b"relative vorticity of fluid", b"1/s"),
             "divergence": (b"divergence of fluid", b"1/s"),
             "absolute": (b"absolute vorticity of fliud layer", b"1/s"),
             "kinetic_energy": (b"kinetic energy of fluid", b"1/s"),
             "tracer_1": (b"tracer field #1", b"tracer per unit thickness"),
             "tracer_2": (b"tracer field #2", b"1/s"),
             "tracer_3": (b"tracer field #3", b"tracer per unit thickness"),
             "tracer_4": (b"tracer field #4", b"tracer per unit thickness"),
             "tracer_5": (b"tracer field #5", b"tracer per unit thickness")}
    for k, (d, u) in descr.items():
        assert v[k].dimensions == ("time", "grid_cells") and v[k].data.dtype.str == ">f4", k
        at = dict(v[k]._attributes)
        assert list(at) == ["long_name", "title", "units", "alt_units", "type", "positions", "missing_value",
                            "mult_factor", "contour_min", "contour_int"], k
        assert (at["long_name"], at["title"], at["units"], at["alt_units"], at["type"], at["positions"]) == \
            (d, d, u, b"no qp alternate units", b"xxx", b"center"), k
        for key, val in (("missing_value", 1e36), ("mult_factor", 1.0), ("contour_min", 1e36),
                         ("contour_int", 1e36)):
            assert at[key].dtype == np.float32 and at[key] == np.float32(val), (k, key)
    # grid_dims is defined and never written: the netCDF int fill value
    assert v["grid_dims"][:].tolist() == [-2147483647]
    nc.close()
    q.close()


def _calendar(secs):
    year = secs // (365 * 86400)
    day = datetime.date(2001, 1, 1) + datetime.timedelta(days=int(secs % (365 * 86400) // 86400))
    return day.strftime("%m/%d/") + "%04d" % year


def test_qp_records_fields_and_times(grid3, tmp_path):
    g = grid3
    rng = np.random.default_rng(7)
    s = (g.iim, g.jjm, 10)
    t0 = 86400 * 30
    q = S.QpFile(g, "run with trailing blanks   ", t0, str(tmp_path / "qp.nc"))
    times = [t0, t0 + 86400 * 29 + 3601, 86400 * (365 * 3 + 58) + 17, 86400 * (365 * 12 + 364) + 86399]
    fields = []
    for r, t in enumerate(times):
        f = {}
        for k, name in enumerate(S.QP_FIELDS):
            if (k + r) % 3:   # some fields left unpacked in each record
                f[name] = rng.standard_normal(s) * 10.0 ** rng.integers(-6, 6)
                q.pack_field(name, f[name])
        q.write(t)
        fields.append(f)
    assert q.records == len(times)
    nc = _read(tmp_path / "qp.nc")
    assert nc._attributes["run"] == b"run with trailing blanks"
    v = nc.variables
    for r, t in enumerate(times):
        start = times[r - 1] if r else t0
        mid = (start + t) // 2
        assert v["start_period"][r] == start / 86400.0 and v["end_period"][r] == t / 86400.0
        assert v["period_length"][r] == (t - start) / 86400.0
        assert v["time"][r] == mid / 86400.0
        assert b"".join(v["char_time"][r]).decode() == _calendar(mid)
        for name in S.QP_FIELDS:
            got = v[name][r]
            if name in fields[r]:
                assert np.array_equal(got, _packed(g, fields[r][name]).astype(np.float32)), (r, name)
            else:
                assert not got.any(), (r, name)
    nc.close()
    # a record's time strings at known dates of the no-leap calendar
    assert [_calendar(d * 86400) for d in (0, 31, 58, 59, 364, 365)] == \
        ["01/01/0000", "02/01/0000", "02/28/0000", "03/01/0000", "12/31/0000", "01/01/0001"]
    q.close()


def test_qp_coordinates(tmp_path):
    for level in (2, 4):
        g = S.Grid(level)
        path = tmp_path / f"qp{level}.nc"
        S.QpFile(g, "SWTC#5", 0, str(path)).close()
        nc = _read(path)
        v = nc.variables
        lat, lon = v["grid_center_lat"][:].copy(), v["grid_center_lon"][:].copy()
        clat, clon = v["grid_corner_lat"][:].copy(), v["grid_corner_lon"][:].copy()
        idx = v["interp_indx"][:].astype(np.int64)
        nc.close()
        n = 10 * 4 ** level + 2
        assert sorted(idx.tolist()) == list(range(1, n + 1))
        # the centres: the generated points of the global cells, in the cells'
        # packed order (panel, j, i) of the real slots
        p_list = g.array("p_list").reshape(3, -1, order="F").T
        P = _xyz(lat, lon)
        assert np.abs(P - p_list[idx - 1]).max() < 1e-14
        gi = g.array("grid_center_index")
        assert np.array_equal(idx, _packed(g, gi))
        # the corners: each is one of the grid's corners adjacent to the cell
        c = g.array("c").reshape(3, -1, order="F").T
        cmask = (g.array("grid_corner_mask") > 0.5).ravel(order="F")
        creal = c[cmask]
        cll = _xyz(np.arcsin(creal[:, 2]), np.arctan2(creal[:, 1], creal[:, 0]))
        written = _xyz(clat, clon).reshape(-1, 3)
        d, which = cKDTree(cll).query(written)
        assert d.max() < 1e-12
        _, near = cKDTree(p_list[idx - 1]).query(creal[which], k=3)
        cell = np.repeat(np.arange(n), 6)
        assert np.all((near == cell[:, None]).any(axis=1))
        # counter-clockwise about the outward normal; 12 pentagons repeat
        # their fifth corner; the polygons tile the sphere
        C = _xyz(clat, clon)
        pent = np.all(C[:, 5] == C[:, 4], axis=1)
        assert pent.sum() == 12
        area = 0.0
        for k in range(6):
            a, b = C[:, k], C[:, (k + 1) % 6]
            same = np.all(a == b, axis=1)
            assert np.array_equal(same, pent if k == 4 else np.zeros(n, bool)), k
            det = np.einsum("ij,ij->i", P, np.cross(a - P, b - P))
            assert det[~same].min() > 0, k
            num = np.einsum("ij,ij->i", P, np.cross(a, b))
            den = 1 + np.einsum("ij,ij->i", P, a) + np.einsum("ij,ij->i", P, b) + np.einsum("ij,ij->i", a, b)
            area += np.where(same, 0.0, 2 * np.arctan2(num, den)).sum()
        assert abs(area - 4 * np.pi) < 1e-11


def _blank_history(raw):
    k = raw.index(b"Created: ")
    return raw[:k] + b" " * 39 + raw[k + 39:]


def test_qp_caller_coordinates(grid3, tmp_path):
    g = grid3
    p = g.array("p")
    c = g.array("c")
    coords = (np.arcsin(p[2]), np.arctan2(p[1], p[0]), np.arcsin(c[2]), np.arctan2(c[1], c[0]))
    S.QpFile(g, "SWTC#5", 0, str(tmp_path / "a.nc")).close()
    S.QpFile(g, "SWTC#5", 0, str(tmp_path / "b.nc"), coords=coords).close()
    a, b = (tmp_path / "a.nc").read_bytes(), (tmp_path / "b.nc").read_bytes()
    assert len(a) == len(b)
    # numpy's arcsin/arctan2 and the C library's may differ in the last bit
    na, nb = _read(tmp_path / "a.nc"), _read(tmp_path / "b.nc")
    for k in ("grid_center_lat", "grid_center_lon", "grid_corner_lat", "grid_corner_lon"):
        assert np.abs(na.variables[k][:] - nb.variables[k][:]).max() <= 1e-15, k
    na.close()
    nb.close()
    shifted = (coords[0] * 0.5,) + coords[1:]
    S.QpFile(g, "SWTC#5", 0, str(tmp_path / "c.nc"), coords=shifted).close()
    nc = _read(tmp_path / "c.nc")
    assert np.array_equal(nc.variables["grid_center_lat"][:], _packed(g, shifted[0]))
    nc.close()
    # the header and everything but the coordinates are identical
    hdr = len(a) - 116 * g.max_ig - 4   # the non-record data: 116 bytes a cell and grid_dims
    assert _blank_history(a)[:hdr] == _blank_history(b)[:hdr]


def test_qp_name_and_errors(grid3, tmp_path):
    assert S.qp_name(0) == "qp_output/hswm_d000000p000.g2.nc"
    assert S.qp_name(86400 - 1) == "qp_output/hswm_d000000p000.g2.nc"
    assert S.qp_name(86400 * 12345 + 5, 7) == "qp_output/hswm_d012345p007.g2.nc"
    assert S.qp_name(86400, 123) == "qp_output/hswm_d000001p123.g2.nc"
    q = S.QpFile(grid3, "x", 0, str(tmp_path / "qp.nc"))
    with pytest.raises(S.SwmError):
        q.pack_field("pressure", np.zeros((grid3.iim, grid3.jjm, 10)))
    with pytest.raises(S.SwmError):
        q.write(-1)
    os.remove(tmp_path / "qp.nc")
    with pytest.raises(S.SwmError):
        q.write(0)
    assert q.records == 0
    q.close()
    with pytest.raises(S.SwmError):
        S.QpFile(grid3, "x", 0, str(tmp_path / "no_such_dir" / "qp.nc"))
