"""GPU: the device time loop captured as a HIP graph (step.CapturedSteps)
equals the eager loop bit for bit on every replay, with eager work and
synchronizations between the replays, for the single-patch split step (each
split mode) and the batched ensemble step."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dclaw_amd import _lib
    _lib.lib()
    return torch


def _eager_work(torch):
    x = torch.randn(4096, device="cuda", dtype=torch.float64)
    float((x * x).sum().item())
    torch.cuda.synchronize()


@pytest.mark.parametrize("mode", ["ordered", "single", "dry"])
def test_captured_patch_loop_equals_eager(dev, mode):
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    torch = dev
    pt = W.fractal_sources(260, 180, dx=2.0, seed=2, nsrc=5)
    print(f"[graph] patch 260x180 mode {mode}", flush=True)
    e = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy, split_mode=mode)
    g = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy, split_mode=mode)
    for r in (e, g):
        r.set_dt(1e-3)
        r.cfl.zero_()
    cap = S.CapturedSteps(g, 6, 0.9, 1e9)
    for rep in range(4):
        for _ in range(6):
            e.step()
            S.dt_update(e.dt, e.cfl, 0.9, 1e9)
        cap.replay()
        _eager_work(torch)
        assert np.array_equal(g.q(), e.q()), rep
        assert torch.equal(g.dt, e.dt), rep
    assert np.abs(e.q()[0]).max() > 0.0


def test_captured_ensemble_loop_equals_eager(dev):
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    torch = dev
    members = W.ensemble(6, 60, 220, seed=5)
    e = S.Ensemble(members, seg_rows=16)
    g = S.Ensemble(members, seg_rows=16)
    for r in (e, g):
        r.dt.fill_(1e-3)
        r.cfl.zero_()
    cap = S.CapturedSteps(g, 4, 0.9, 1e9)
    for rep in range(3):
        for _ in range(4):
            e.step()
            S.dt_update(e.dt, e.cfl, 0.9, 1e9)
        _eager_work(torch)
        cap.replay()
        torch.cuda.synchronize()
        for k in range(len(members)):
            assert np.array_equal(g.q(k), e.q(k)), (rep, k)
        assert torch.equal(g.dt, e.dt), rep
