"""Restart files (swm/source/swm_restart.F over utilities/layer_2/io_utilities.F)
against the compiled reference, on the CPU: the library's file codec and the
oracle.

restart_g3.npz (tests/golden/make_golden.py) holds, at g = 3 over a mountain:
  a_*  a file the reference wrote after 3 steps, and its state then;
  b_*  the reference's continuation from that file (read_swm_restart, then a
       forward step and ABM3);
  c_*  the reference's continuation from a file the library wrote, and the
       file the reference wrote after it.
Bars: bit for bit on the file bytes the reference defines and on every value
read; the oracle (pinned bit for bit to the reference step) continues from
what the library reads exactly as the reference did.
"""
import os

import numpy as np
import pytest

import swm_mi355x as S
from oracle.oracle import Oracle, grid_arrays_from

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NT = 5
HEAD_NAME, HEAD_DIMS = 40, 4


@pytest.fixture(scope="module")
def ref():
    return np.load(os.path.join(GOLD, "restart_g3.npz"))


@pytest.fixture(scope="module")
def grid():
    return S.Grid(3)


def _write(tmp_path, name, data):
    p = tmp_path / name
    p.write_bytes(bytes(data))
    return str(p)


def _defined_bytes(raw):
    """The file with the bytes the reference leaves undefined zeroed: the
    var_dims entries beyond nvar_dims in the header records (io_utilities.F
    pack_header sets only those it uses)."""
    b = bytearray(raw)
    nhead, lrecl = np.frombuffer(raw[:8], "<i4")
    nvar = int(np.frombuffer(raw[lrecl:lrecl + 4], "<i4")[0])
    for v in range(nvar):
        o = (2 + v) * lrecl
        nd = int(np.frombuffer(raw[o + HEAD_NAME:o + HEAD_NAME + 4], "<i4")[0])
        for k in range(nd, HEAD_DIMS):
            q = o + HEAD_NAME + 4 + 4 * k
            b[q:q + 4] = b"\0\0\0\0"
    return bytes(b)


def test_restart_names(ref):
    # swm_restart.F:122-145; the reference's own names for 3 x 800 s and for
    # 2 d 05:01:01 + 2 x 800 s
    assert S.restart_name(2400) == str(ref["a_name"])
    assert S.restart_name(int(ref["c_ts"]) + 1600) == str(ref["c_name"])
    assert S.restart_name(0) == "restarts/swm_restart.d000000.00:00:00"
    assert S.restart_name(86400 * 123456 + 3600 * 23 + 60 * 59 + 59) == "restarts/swm_restart.d123456.23:59:59"
    with pytest.raises(S.SwmError):
        S.restart_name(-1)


def test_file_layout_of_reference(ref):
    # the facts the codec relies on, read off the reference's file
    raw = bytes(ref["a_file"])
    nhead, lrecl = np.frombuffer(raw[:8], "<i4")
    assert (nhead, lrecl) == (6, (5 + NT) * 8)
    assert list(np.frombuffer(raw[lrecl:lrecl + 16], "<i4")) == [3, 8, 642, 1]
    assert len(raw) == (nhead + 642) * lrecl
    assert int(np.frombuffer(raw[(nhead - 1) * lrecl:(nhead - 1) * lrecl + 8], "<i8")[0]) == 2400


def test_reader_recovers_reference_state(ref, grid, tmp_path):
    ts, h, u, tr = S.read_restart_file(grid, _write(tmp_path, "a", ref["a_file"]), NT)
    assert ts == 2400
    # h and tracers: every slot (halos hold their global cell's value, as the
    # reference's wrap leaves them); u: the real cells (the halo corners of
    # the reference's state are rotated by its vrtx wrap)
    assert np.array_equal(h.reshape(-1, order="F"), ref["a_s3_h"])
    assert np.array_equal(tr.reshape(-1, order="F"), ref["a_s3_tracer"])
    m = grid.array("grid_center_mask") == 1.0
    ua = ref["a_s3_u"].reshape(u.shape, order="F")
    assert np.array_equal(u[:, :, m], ua[:, :, m])
    # the reference's own file written after its continuation from ours
    ts, h, u, tr = S.read_restart_file(grid, _write(tmp_path, "c", ref["c_file"]), NT)
    assert ts == int(ref["c_ts"]) + 1600
    assert np.array_equal(h.reshape(-1, order="F"), ref["c_s2_h"])
    assert np.array_equal(tr.reshape(-1, order="F"), ref["c_s2_tracer"])
    uc = ref["c_s2_u"].reshape(u.shape, order="F")
    assert np.array_equal(u[:, :, m], uc[:, :, m])


def test_writer_matches_reference_bytes(ref, grid, tmp_path):
    s = (grid.iim, grid.jjm, 10)
    p = str(tmp_path / "mine")
    S.write_restart_file(grid, p, ref["a_s3_h"].reshape(s, order="F"),
                         ref["a_s3_u"].reshape((2, 2) + s, order="F"),
                         ref["a_s3_tracer"].reshape((grid.iim, grid.jjm, NT, 10), order="F"), 2400)
    mine = open(p, "rb").read()
    theirs = bytes(ref["a_file"])
    assert len(mine) == len(theirs)
    assert _defined_bytes(mine) == _defined_bytes(theirs)
    # and the file the reference wrote after reading ours
    p = str(tmp_path / "mine_c")
    S.write_restart_file(grid, p, ref["c_s2_h"].reshape(s, order="F"), ref["c_s2_u"].reshape((2, 2) + s, order="F"),
                         ref["c_s2_tracer"].reshape((grid.iim, grid.jjm, NT, 10), order="F"), int(ref["c_ts"]) + 1600)
    assert _defined_bytes(open(p, "rb").read()) == _defined_bytes(bytes(ref["c_file"]))


def _fields_equal(orc, ref, prefix):
    for name in ("h", "eta", "div", "psi", "chi", "ke", "relative", "u", "tracer"):
        mine = orc.field(name).reshape(-1, order="F")
        assert np.array_equal(mine, ref[f"{prefix}_{name}"]), f"{prefix} {name}"


def test_reference_continuation_from_its_file(ref, grid, tmp_path):
    # read_swm_restart, then 4 steps (the first a forward step): the oracle
    # from the arrays the library's reader returns, with the reference's hs
    _, h, u, tr = S.read_restart_file(grid, _write(tmp_path, "a", ref["a_file"]), NT)
    orc = Oracle(grid_arrays_from(grid), NT, float(ref["dt"]), S.REF_GRAV, hs=ref["ic_hs"])
    orc.init_state(h, u, tr)
    _fields_equal(orc, ref, "b_s0")
    orc.step(1)
    _fields_equal(orc, ref, "b_s1")
    orc.step(3)
    _fields_equal(orc, ref, "b_s4")


def test_reference_continuation_from_our_file(ref, grid, tmp_path):
    # the reference read a file of the library's writer: the state it started
    # from is the writer's input at the real cells, and the oracle continues
    # from what the library's reader returns exactly as the reference did
    p = str(tmp_path / "ours")
    S.write_restart_file(grid, p, ref["c_in_h"], ref["c_in_u"], ref["c_in_tracer"], int(ref["c_ts"]))
    ts, h, u, tr = S.read_restart_file(grid, p, NT)
    assert ts == int(ref["c_ts"])
    m = grid.array("grid_center_mask") == 1.0
    assert np.array_equal(h[m], ref["c_in_h"][m])
    assert np.array_equal(u[:, :, m], ref["c_in_u"][:, :, m])
    assert np.array_equal(tr.transpose(0, 1, 3, 2)[m], ref["c_in_tracer"].transpose(0, 1, 3, 2)[m])
    orc = Oracle(grid_arrays_from(grid), NT, float(ref["dt"]), S.REF_GRAV, hs=ref["ic_hs"])
    orc.init_state(h, u, tr)
    _fields_equal(orc, ref, "c_s0")
    orc.step(2)
    _fields_equal(orc, ref, "c_s2")


def test_reader_fp32_and_plain_files(ref, grid, tmp_path):
    # io_utilities.F reads real_bytes = 4 files and files without a time
    # stamp record (restart flag false): build both from the reference's file
    raw = bytes(ref["a_file"])
    nhead, lrecl = (int(x) for x in np.frombuffer(raw[:8], "<i4"))
    ncell = 642
    nval = 5 + NT
    data = np.frombuffer(raw[nhead * lrecl:], "<f8").reshape(ncell, lrecl // 8)[:, :nval]
    # FP32 restart: lrecl = max(nval * 4, 64)
    l4 = max(nval * 4, 64)
    out = bytearray((nhead + ncell) * l4)
    out[0:8] = np.array([nhead, l4], "<i4").tobytes()
    out[l4:l4 + 16] = np.array([3, 4, ncell, 1], "<i4").tobytes()
    for v in range(3):
        out[(2 + v) * l4:(2 + v) * l4 + 64] = raw[(2 + v) * lrecl:(2 + v) * lrecl + 64]
    out[(nhead - 1) * l4:(nhead - 1) * l4 + 8] = raw[(nhead - 1) * lrecl:(nhead - 1) * lrecl + 8]
    for k in range(ncell):
        o = (nhead + k) * l4
        out[o:o + 4 * nval] = data[k].astype("<f4").tobytes()
    ts, h, u, tr = S.read_restart_file(grid, _write(tmp_path, "f4", out), NT)
    assert ts == 2400
    assert np.array_equal(h.reshape(-1, order="F"), ref["a_s3_h"].astype(np.float32).astype(np.float64))
    # no time stamp record: nhead = nvar + 2, restart flag 0
    plain = bytearray(raw[:lrecl]) + bytearray(raw[lrecl:(nhead - 1) * lrecl]) + bytearray(raw[nhead * lrecl:])
    plain[0:4] = np.array([nhead - 1], "<i4").tobytes()
    plain[lrecl + 12:lrecl + 16] = np.array([0], "<i4").tobytes()
    ts, h2, _, _ = S.read_restart_file(grid, _write(tmp_path, "plain", plain), NT)
    assert ts == 0
    assert np.array_equal(h2.reshape(-1, order="F"), ref["a_s3_h"])


def test_reader_refuses_bad_files(ref, grid, tmp_path):
    good = _write(tmp_path, "a", ref["a_file"])
    with pytest.raises(S.SwmError) as e:
        S.read_restart_file(grid, good, NT - 1)   # tracer count
    assert e.value.status == S.SWM_ERR_DIMS
    with pytest.raises(S.SwmError) as e:
        S.read_restart_file(S.Grid(2), good, NT)   # another grid
    assert e.value.status == S.SWM_ERR_DIMS
    with pytest.raises(S.SwmError):
        S.read_restart_file(grid, str(tmp_path / "missing"), NT)
    with pytest.raises(S.SwmError):
        S.read_restart_file(grid, _write(tmp_path, "short", bytes(ref["a_file"])[:-100]), NT)
    raw = bytearray(ref["a_file"])
    lrecl = int(np.frombuffer(bytes(raw[4:8]), "<i4")[0])
    raw[2 * lrecl:2 * lrecl + 1] = b"X"   # "H" renamed: variable missing
    with pytest.raises(S.SwmError):
        S.read_restart_file(grid, _write(tmp_path, "noh", raw), NT)
