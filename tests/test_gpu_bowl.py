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


def test_bowl_ensemble_equals_oracle(dev, oracle_mod):
    """Bowls of four sizes stepped together by the batched step equal the
    oracle's per-patch loop bit for bit, each with its own dt."""
    from dclaw_amd import step as S
    O = oracle_mod
    members = [W.Member(W.bowl_slosh(n), seed=k) for k, n in enumerate((41, 64, 97, 130))]
    nsteps = 30
    print("[bowl] ensemble sizes", [m.patch.mx for m in members], flush=True)
    ens = S.Ensemble(members, seg_rows=16)
    ens.run(nsteps, 1e-4, 0.75, 1e99)
    dev.cuda.synchronize()
    I = (slice(None), slice(2, -2), slice(2, -2))
    for k, m in enumerate(members):
        pt = m.patch
        qo, _, dto, _, _ = O.run(O.params_from(pt.params), pt.q, pt.aux, pt.dx, pt.dy, nsteps,
                                 1e-4, 0.75, 1e99)
        assert np.array_equal(ens.q(k)[I], qo[I]), k
        assert ens.dt[k].item() == pytest.approx(dto, rel=1e-14), k


def test_bowl_adjoint_is_transpose_of_tangent(dev):
    """<lambda, TL(dq, dP)> = <J^T lambda, dq> + g . a at a sloshing bowl
    state (moving shoreline over a curved bed), GPU tangent-linear step
    against GPU adjoint step."""
    from dclaw_amd import step as S
    from dclaw_amd import tangent as T
    from dclaw_amd.params import CParams, to_c
    from test_gpu_adjoint import _gpu_adjoint
    from test_emu_adjoint import _scaled
    torch = dev
    pt = W.bowl_slosh(64)
    print("[bowl] adjoint 64x64", flush=True)
    r = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy)
    r.run(20, 1e-4, 0.75, 1e99)
    q = r.q()
    dt = float(r.dt.item())
    rng = np.random.default_rng(5)
    lam = np.asfortranarray(rng.normal(size=q.shape))
    for m in range(7):
        lam[m] /= np.abs(q[m]).max() + 1.0
    names = ["phi", "kref", "mu"]
    c = to_c(pt.params)
    scales = [abs(getattr(c, n)) for n in names]
    dPs = [CParams(**{n: s}) for n, s in zip(names, scales)]
    bc = (1, 1, 1, 1)
    qo, lin, g = _gpu_adjoint(torch, pt, q, np.asfortranarray(pt.aux), dt, bc, lam, dPs)
    dq = _scaled(rng, q)
    a = rng.normal(size=len(names))
    dP = T.param_direction(**{n: ak * s for n, ak, s in zip(names, a, scales)})
    qi, dqi = S.to_device_q(q), S.to_device_q(dq)
    qt, dqt = torch.zeros_like(qi), torch.zeros_like(qi)
    cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
    dtd = torch.tensor([dt], dtype=torch.float64, device="cuda")
    T.stepgrid_tl(qi, dqi, qt, dqt, S.bed_tensor(pt.aux), 1, pt.mx, pt.my, pt.dx, pt.dy, bc, dtd,
                  S.params_tensor([pt.params]), T._record_tensor(dP), cfl)
    torch.cuda.synchronize()
    assert np.array_equal(S.to_host_q(qt), qo)
    tl = S.to_host_q(dqt)
    lhs = float((lam * tl).sum())
    rhs = float((lin * dq).sum()) + float(g @ a)
    assert abs(lhs - rhs) <= 1e-12 * float(np.abs(lam * tl).sum()), (lhs, rhs)
