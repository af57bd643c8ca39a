"""The C-ABI library: it loads, exports every entry point include/dclaw_hip.h
declares, and rejects bad arguments before touching a device (no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "dclaw_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*([a-z_][a-z0-9_]*)\s*\(",
                       src, flags=re.M)
    return sorted(set(names))


@pytest.fixture(scope="module")
def L():
    from dclaw_amd import _lib
    return _lib.lib()


def test_every_declared_symbol_is_exported(L):
    names = declared_functions()
    assert "dclaw_stepgrid_dev" in names and "stepgrid_dclaw_" in names
    from dclaw_amd import _lib
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, f"{n} has no ctypes signature"


def test_version_and_width(L):
    assert L.dclaw_version() == 1
    assert L.dclaw_strip_width() == 58


def test_argument_validation(L):
    from dclaw_amd import _lib
    rc = L.dclaw_stepgrid_dev(None, None, None, 1, 10, 10, 1.0, 1.0, None, None, None, None)
    assert rc == -1 and b"null" in L.dclaw_last_error()
    rc = L.dclaw_stepgrid_dev(1, 1, 2, 1, 10, 10, 1.0, 1.0, 3, 4, 5, None)
    assert rc == -1 and b"alias" in L.dclaw_last_error()
    rc = L.dclaw_stepgrid_dev(1, 2, 3, 1, 0, 10, 1.0, 1.0, 3, 4, 5, None)
    assert rc == -1
    bc = (ctypes.c_int * 4)(1, 1, 2, 1)
    assert L.dclaw_bc_fill(1, 10, 10, bc, None) == -1       # periodic is not supported
    mx = (ctypes.c_int * 3)(64, 200, 513)
    my = (ctypes.c_int * 3)(64, 100, 65)
    assert L.dclaw_worklist_size(mx, my, 3, 32) == 2 * 2 + 4 * 4 + 9 * 3
    tiles = (ctypes.c_int32 * (4 * 47))()
    assert L.dclaw_worklist_build(mx, my, 3, 32, tiles) == 47
    assert tiles[0] == 2            # largest patch first
    with pytest.raises(_lib.DClawError):
        _lib.check(L.dclaw_worklist_size(mx, my, 3, 0))


def test_params_struct_layout():
    from dclaw_amd.params import CParams
    from oracle.oracle import OracleParams
    assert [f[0] for f in CParams._fields_] == [f[0] for f in OracleParams._fields_]
    assert ctypes.sizeof(CParams) == 20 * 8 + 14 * 4 + 3 * 8 + 3 * 4 + 4     # + padding


def test_stage_argument_validation(L):
    """The per-stage entry points reject bad arguments before touching the GPU."""
    assert L.dclaw_b4step2_dev(1, 2, 1, 10, 10, 1.0, 1.0, 3, None, None) == -1
    assert b"maux" in L.dclaw_last_error()
    assert L.dclaw_step2_dev(1, 2, 6, 10, 10, 1.0, 1.0, 3, 4, None, 5, 6, 7, 8, None) == -1
    assert b"null" in L.dclaw_last_error()
    assert L.dclaw_step2_dev(1, 2, 6, 10, 10, 0.0, 1.0, 3, 4, 5, 5, 6, 7, 8, None) == -1
    assert L.dclaw_src2_dev(None, 2, 6, 10, 10, 3, 4, None, None) == -1
    assert L.dclaw_check4nans_dev(1, 0, 10, 2, None) == -1
    assert L.dclaw_check4nans_dev(1, 100000, 100000, 2, None) == -1   # flat index overflows


def test_split_workspace_and_dry_rows(L):
    """Host-side sizing of the split step's workspace and the dry-pass
    setting: 2 ints per tile plus the two list counters; bad values rejected
    without changing the setting."""
    prev = L.dclaw_set_seg_rows(12)
    try:
        nt = L.dclaw_tile_count(4096, 4096)
        assert nt == 71 * 342
        assert L.dclaw_split_workspace_size(4096, 4096) == 3 * nt + 4
        assert L.dclaw_split_workspace_size(1, 1) == 7
        assert L.dclaw_split_workspace_size(0, 5) == -1
    finally:
        L.dclaw_set_seg_rows(prev)
    m = L.dclaw_set_split_mode(1)
    assert m in (0, 1, 2)
    assert L.dclaw_set_split_mode(4) == -1 and b"split mode" in L.dclaw_last_error()
    assert L.dclaw_set_split_mode(m) == 1
    p = L.dclaw_set_dry_rows(96)
    assert p >= 1
    assert L.dclaw_set_dry_rows(0) == -1 and b"dry rows" in L.dclaw_last_error()
    assert L.dclaw_set_dry_rows(p) == 96
    # the split step rejects a workspace smaller than one int per tile
    rc = L.dclaw_stepgrid_split_dev(1, 2, 3, 1, 100, 100, 1.0, 1.0, None, 4, 5, 6, 7, 1, None)
    assert rc == -1 and b"tile_flags" in L.dclaw_last_error()


def test_adjoint_workspace_and_validation(L):
    """The adjoint step's workspace size and its argument checks (host side,
    before any device work)."""
    from dclaw_amd import adjoint as AD
    n = L.dclaw_adjoint_workspace_doubles(10, 6)
    assert n == AD.workspace_doubles(10, 6) and n > 100 * 14 * 10
    assert L.dclaw_adjoint_workspace_doubles(0, 6) == 0
    bc = (ctypes.c_int * 4)(1, 1, 1, 1)
    args = lambda np_, work, nwork, mx=10: (1, 2, 1, mx, 6, 1.0, 1.0, bc, 3, 4, 5, np_, 6, 7, 8,
                                            9, work, nwork, None)
    assert L.dclaw_adjoint_step_dev(*args(5, 10, n)) == -1
    assert b"np" in L.dclaw_last_error()
    assert L.dclaw_adjoint_step_dev(*args(1, 10, n - 1)) == -1
    assert b"workspace" in L.dclaw_last_error()
    assert L.dclaw_adjoint_step_dev(*args(1, None, n)) == -1
    assert L.dclaw_adjoint_step_dev(*args(1, 10, n, mx=1)) == -1
    assert b"mx, my >= 2" in L.dclaw_last_error()
    assert L.dclaw_gauges_adj_dev(1, 2, 1, 10, 6, 0.0, 0.0, 1.0, 1.0, 1e-3, 3, 0, 4, 5, None) == -1


def test_handle_settings(L):
    """Handles take their settings from the library-wide values at creation,
    keep their own afterwards and reject bad values; no device memory is
    touched before a step."""
    import ctypes
    prev = L.dclaw_set_seg_rows(12)
    try:
        h = ctypes.c_void_p()
        assert L.dclaw_create(ctypes.byref(h)) == 0 and h.value
        assert L.dclaw_handle_tile_count(h, 300, 200) == 6 * 17
        assert L.dclaw_handle_set_seg_rows(h, 4) == 12
        assert L.dclaw_set_seg_rows(12) == 12          # the library-wide value is unchanged
        assert L.dclaw_handle_tile_count(h, 300, 200) == 6 * 50
        assert L.dclaw_handle_set_seg_rows(h, 0) == -1 and b"seg_rows" in L.dclaw_last_error()
        assert L.dclaw_handle_set_split_mode(h, 3) == 1
        assert L.dclaw_handle_set_split_mode(h, 4) == -1
        assert L.dclaw_handle_set_dry_rows(h, 16) == 48
        assert L.dclaw_handle_tile_count(h, 0, 5) == -1
        assert L.dclaw_handle_state_written(h, None) == 0     # nothing allocated yet
        assert L.dclaw_handle_stepgrid_dev(None, 1, 2, 3, 1, 10, 10, 1.0, 1.0, None, 4, 5, 6,
                                           None) == -1
        assert b"null handle" in L.dclaw_last_error()
        assert L.dclaw_destroy(h) == 0
        assert L.dclaw_destroy(None) == 0
        assert L.dclaw_handle_set_seg_rows(None, 4) == -1
        assert L.dclaw_create(None) == -1
    finally:
        L.dclaw_set_seg_rows(prev)
