"""GPU tests of the device-side plumbing around the fused step: ghost fill,
device dt rule and short chained runs, each against the oracle."""
import numpy as np
import pytest

from helpers import TOL_STEP, mask_equal, rel_linf_wet, slope_patch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dclaw_amd import _lib
    _lib.lib()
    return torch


@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (3, 3, 3, 3), (1, 3, 3, 1)])
def test_bc_fill_matches_oracle(dev, oracle_mod, bc):
    from dclaw_amd import step as S
    O = oracle_mod
    rng = np.random.default_rng(3)
    q = np.asfortranarray(rng.standard_normal((7, 37 + 4, 23 + 4)))
    t = S.to_device_q(q)
    S.bc_fill(t, 37, 23, bc)
    dev.cuda.synchronize()
    assert np.array_equal(S.to_host_q(t), O.bc_fill(q, bc))


def test_dt_update(dev):
    from dclaw_amd import step as S
    dt = dev.tensor([1e-3, 2e-3, 5e-3], dtype=dev.float64, device="cuda")
    cfl = dev.tensor([0.5, 0.0, 4.0], dtype=dev.float64, device="cuda")
    S.dt_update(dt, cfl, 0.9, 0.01)
    dev.cuda.synchronize()
    assert dt.cpu().tolist() == [min(1e-3 * 0.9 / 0.5, 0.01), 0.01, min(5e-3 * 0.9 / 4.0, 0.01)]
    assert cfl.cpu().tolist() == [0.0, 0.0, 0.0]


@pytest.mark.parametrize("nsteps", [1, 2, 10])
def test_runner_short(dev, oracle_mod, nsteps):
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    O = oracle_mod
    pt = W.inclined_plane_mound()
    qo, _, dto, cflo, dtso = O.run(O.params_from(pt.params), pt.q, pt.aux, pt.dx, pt.dy, nsteps,
                                   1e-3, 0.9, 1e9)
    r = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy)
    hist = dev.zeros(nsteps, dtype=dev.float64, device="cuda")
    r.run(nsteps, 1e-3, 0.9, 1e9, cfl_hist=hist)
    dev.cuda.synchronize()
    qg = r.q()
    print("cfl gpu", hist.cpu().numpy()[:5], "oracle", cflo[:5], "dt", r.dt.item(), dto)
    err = rel_linf_wet(qg[:, 2:-2, 2:-2], qo[:, 2:-2, 2:-2], pt.params.dry_tolerance)
    print("err", err)
    assert mask_equal(qg[:, 2:-2, 2:-2], qo[:, 2:-2, 2:-2], pt.params.dry_tolerance)
    assert err <= TOL_STEP * 100
