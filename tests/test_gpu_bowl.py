"""GPU: the reference's bowl-slosh case (tests/bowl-slosh: clear water
sloshing in a paraboloid bowl on 41 x 41 cells of [-2, 2]^2, output times 0.5
and 1.0, cfl_desired 0.75, dt_initial 1e-4; setrun.py:70-167) run by the
device-resident loop in every split mode equals the oracle's loop bit for bit,
outputs, step times and gauges included, and stays close to Thacker's exact
solution (docs/src/tests.rst:64-84)."""
import numpy as np
import pytest

from dclaw_amd import workloads as W
from helpers import bowl_errors
from test_gpu_loop import _oracle_run_until

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dclaw_amd import _lib
    _lib.lib()
    return torch


@pytest.mark.parametrize("mode", ["ordered", "single", "dry", "hybrid"])
def test_bowl_slosh_equals_oracle_and_thacker(dev, oracle_mod, mode):
    from dclaw_amd import step as S
    O = oracle_mod
    n = 41
    pt = W.bowl_slosh(n)
    outs = [0.5, 1.0]
    # gauges at the bowl's centre and on the initial shoreline (x from the
    # domain's lower edge)
    xy = np.array([[2.0, 2.0], [2.5, 2.0], [3.45, 2.0], [2.5, 2.95]])
    print(f"[bowl] {n}x{n} mode {mode} outputs {outs}", flush=True)
    r = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy, split_mode=mode)
    res = r.run_until(outs, 1e-4, 0.75, 1e99, gauges_xy=xy, check_nans=True)
    qo, reached, times, gv = _oracle_run_until(O, pt, outs, 1e-4, 0.75, 1e99, xy)
    assert res.outputs == reached
    assert np.array_equal(res.times, times)
    assert np.array_equal(res.gauge_values, gv)
    I = (slice(None), slice(2, -2), slice(2, -2))
    q = r.q()
    assert np.array_equal(q[I], qo[I])
    eh, ev = bowl_errors(q[I], n, res.outputs[-1])
    assert eh < 1e-3 and ev < 0.25, (eh, ev)
    loss = q[I][0].sum() / pt.q[I][0].sum() - 1
    assert -0.03 < loss <= 1e-12, loss
