"""GPU: dclaw_stepgrid_split_static_dev, the split step whose tile-flag
workspace is the library's static device array (no allocation, no caller
workspace), against dclaw_stepgrid_split_dev with a caller workspace, on two
calls in a row, the second on a smaller patch (the static array then holds
the first call's flags), and the rejection of a patch with too many tiles."""
import ctypes

import pytest

from helpers import advance_oracle, slope_patch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dclaw_amd import _lib
    _lib.lib()
    return torch


def _case(dev, oracle_mod, mx, my, seed):
    """(inputs on the device, bc) of a flowing patch after a few oracle steps."""
    from dclaw_amd import step as S
    pt = slope_patch(mx, my, slope=0.8, phi_deg=24.0, seed=seed)
    q, aux, dt = advance_oracle(oracle_mod, pt, 4)
    return pt, S.to_device_q(q), S.bed_tensor(aux), dt


def test_static_workspace_step_equals_workspace_step(dev, oracle_mod):
    from dclaw_amd import _lib
    from dclaw_amd import step as S
    L = _lib.lib()
    torch = dev
    bc = (ctypes.c_int * 4)(1, 3, 1, 1)
    for mx, my, seed in [(300, 200, 1), (120, 90, 2)]:
        pt, qi, bed, dt = _case(dev, oracle_mod, mx, my, seed)
        print(f"[no-alloc] {mx}x{my} seed {seed}", flush=True)
        prm = S.params_tensor([pt.params])
        dtd = torch.tensor([dt], dtype=torch.float64, device="cuda")
        nws = L.dclaw_split_workspace_size(mx, my)
        assert 0 < nws <= L.dclaw_static_workspace_ints()
        flags = torch.zeros(nws, dtype=torch.int32, device="cuda")
        q_ref, q_st = torch.full_like(qi, float("nan")), torch.full_like(qi, float("nan"))
        c_ref = torch.zeros(1, dtype=torch.float64, device="cuda")
        c_st = torch.zeros(1, dtype=torch.float64, device="cuda")
        # two steps each (the second reads the flags the first wrote)
        src_ref, src_st = qi.clone(), qi.clone()
        for _ in range(2):
            _lib.check(L.dclaw_stepgrid_split_dev(
                S._ptr(src_ref), S._ptr(q_ref), S._ptr(bed), 1, mx, my, pt.dx, pt.dy, bc,
                S._ptr(dtd), S._ptr(prm), S._ptr(c_ref), S._ptr(flags), nws, S._stream()))
            _lib.check(L.dclaw_stepgrid_split_static_dev(
                S._ptr(src_st), S._ptr(q_st), S._ptr(bed), 1, mx, my, pt.dx, pt.dy, bc,
                S._ptr(dtd), S._ptr(prm), S._ptr(c_st), S._stream()))
            torch.cuda.synchronize()
            I = (slice(2, -2), slice(2, -2))
            assert torch.equal(q_st[I], q_ref[I])
            assert torch.equal(c_st, c_ref)
            src_ref, q_ref = q_ref, src_ref
            src_st, q_st = q_st, src_st


def test_static_workspace_rejects_too_many_tiles(dev):
    from dclaw_amd import _lib
    L = _lib.lib()
    prev = L.dclaw_set_seg_rows(1)
    try:
        mx, my = 20000, 4000
        assert L.dclaw_split_workspace_size(mx, my) > L.dclaw_static_workspace_ints()
        rc = L.dclaw_stepgrid_split_static_dev(None, None, None, 1, mx, my, 1.0, 1.0, None, None,
                                               None, None, None)
        assert rc == -1
        assert b"static workspace" in L.dclaw_last_error()
    finally:
        L.dclaw_set_seg_rows(prev)
