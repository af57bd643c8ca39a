"""GPU: handle-based stepping.  Loops with their own library handles (each
its own segment height, split mode and resident workspace) step two patches
interleaved, through a tuner run on one handle, and give bit for bit the
states and dt of the same loops stepped with the fused step and the
library-wide settings, which the handles leave untouched."""
import numpy as np
import pytest

from helpers import slope_patch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dclaw_amd import _lib
    _lib.lib()
    return torch


def test_handles_side_by_side(dev):
    from dclaw_amd import _lib, autotune
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    L = _lib.lib()
    pa = W.fractal_sources(300, 200, dx=2.0, seed=5, nsrc=6)
    pb = slope_patch(130, 70, slope=0.8, phi_deg=25.0, seed=200)
    prev = L.dclaw_set_seg_rows(32)
    prev_mode = L.dclaw_set_split_mode(1)
    try:
        ha = S.Handle(seg_rows=4, split_mode="ordered")
        hb = S.Handle(seg_rows=12, split_mode="hybrid")
        assert ha.tile_count(300, 200) == L.dclaw_handle_tile_count(ha._h, 300, 200)
        assert ha.tile_count(300, 200) == 6 * 50 and hb.tile_count(130, 70) == 3 * 6
        bcs = ((1, 1, 1, 1), (3, 1, 3, 3))
        runs = []
        for p, h, bc in ((pa, ha, bcs[0]), (pb, hb, bcs[1])):
            rh = S.PatchRunner(p.q, p.aux, p.params, p.dx, p.dy, bc=bc, handle=h)
            rf = S.PatchRunner(p.q, p.aux, p.params, p.dx, p.dy, bc=bc, split_src2=False)
            for r in (rh, rf):
                r.set_dt(1e-3)
            runs.append((rh, rf))
        print("[handle] 300x200 seg 4 ordered + 130x70 seg 12 hybrid, 60 steps", flush=True)
        for n in range(60):
            for rh, rf in runs:
                rh.step()
                rf.step()
            if n == 30:
                # the tuner sets the handle's height and mode and restarts its
                # bookkeeping; the library-wide settings stay as they were
                h, _ = autotune.tune_patch(runs[0][0], candidates=(4, 8), reps=2,
                                           split_modes=("ordered", "single"))
                assert h in (4, 8)
                assert L.dclaw_set_seg_rows(32) == 32 and L.dclaw_set_split_mode(1) == 1
        for rh, rf in runs:
            rh.run(40, rh.dt.clone(), 0.9, 1e9)
            rf.run(40, rf.dt.clone(), 0.9, 1e9)
        dev.cuda.synchronize()
        for k, (rh, rf) in enumerate(runs):
            a, b = rh.q(), rf.q()
            assert np.array_equal(a, b), f"patch {k}: {(a != b).sum()} values differ"
            assert rh.dt.item() == rf.dt.item(), k
            assert (a[0] > 0.0).any()
        # a state written from outside: the handle's bookkeeping starts over
        r = runs[0][0]
        r.qa.copy_(S.to_device_q(pa.q))
        r.state_written()
        f = S.PatchRunner(pa.q, pa.aux, pa.params, pa.dx, pa.dy, bc=bcs[0], split_src2=False)
        for x in (r, f):
            x.set_dt(1e-3)
            x.run(20, 1e-3, 0.9, 1e9)
        dev.cuda.synchronize()
        assert np.array_equal(r.q(), f.q())
        ha.close()
        hb.close()
        ha.close()
    finally:
        L.dclaw_set_seg_rows(prev)
        L.dclaw_set_split_mode(prev_mode)


def test_handles_in_two_threads(dev):
    """Two host threads, each with its own handle (different segment heights
    and split modes) and its own HIP stream, step their patches at the same
    time; each result equals the same loop run alone, bit for bit."""
    import threading
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    pa = W.fractal_sources(300, 200, dx=2.0, seed=7, nsrc=6)
    pb = slope_patch(130, 70, slope=0.8, phi_deg=25.0, seed=201)
    cases = [(pa, (1, 1, 1, 1), 4, "ordered"), (pb, (3, 1, 3, 3), 12, "hybrid")]

    def run(p, bc, seg, mode, out, k):
        s = dev.cuda.Stream()
        with dev.cuda.stream(s):
            h = S.Handle(seg_rows=seg, split_mode=mode)
            r = S.PatchRunner(p.q, p.aux, p.params, p.dx, p.dy, bc=bc, handle=h)
            r.run(50, 1e-3, 0.9, 1e9)
            s.synchronize()
            out[k] = (r.q(), r.dt.item())
            h.close()

    alone = [None, None]
    for k, c in enumerate(cases):
        run(*c, alone, k)
    together = [None, None]
    print("[handle] two threads, 50 steps each", flush=True)
    ts = [threading.Thread(target=run, args=(*c, together, k)) for k, c in enumerate(cases)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    for k in range(2):
        assert together[k] is not None, k
        assert np.array_equal(together[k][0], alone[k][0]), k
        assert together[k][1] == alone[k][1], k
