"""GPU: the batched-adjoint workload at full member size -- 64 members of
1024 x 1024 cells (seeded fractal beds, one Gaussian source each, perturbed
phi, m0, kref), each with its own gauges and gauge-misfit loss.  Every
member's forward run equals the oracle's bit for bit.  The batched
reverse sweep (EnsembleRollout, two checkpoint levels, block checkpoints in
host memory) gives the first and
the last member exactly the gradient of that member's own Rollout, and the
reverse gradient of member 0 is the transpose of its tangent-linear run
(directional derivatives along phi and a depth perturbation agree)."""
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


def _gauges(members, nsteps, seed=11):
    import torch
    rng = np.random.default_rng(seed)
    gxy, tg = [], []
    for m in members:
        # three gauges around the source (the deepest cell), inside the flow
        p = m.patch
        i, j = np.unravel_index(np.argmax(p.q[0][2:-2, 2:-2]), (p.mx, p.my))
        off = rng.uniform(-20.0, 20.0, (3, 2))
        xy = np.column_stack([(i + 0.5 + off[:, 0]) * p.dx, (j + 0.5 + off[:, 1]) * p.dy])
        gxy.append(xy)
        tg.append(torch.tensor(rng.uniform(0.0, 2.0, (nsteps, 3)), device="cuda"))
    return gxy, tg


def test_c4_batched_rollout_gradient(dev):
    from dclaw_amd import adjoint as AD
    from dclaw_amd import step as S
    from dclaw_amd import tangent as T
    from dclaw_amd import workloads as W
    torch = dev
    members = W.ensemble(64, 1024, 1024, seed=1)
    nsteps = 12
    gxy, tg = _gauges(members, nsteps)
    print("[c4] 64 x 1024^2 rollout, 12 steps, checkpoints every 6, sub every 2", flush=True)
    er = AD.EnsembleRollout(members, nsteps, gxy, parameters=("phi", "kref"), checkpoint_every=6,
                            sub_every=2, seg_rows=32, host_checkpoints=True)
    series, qf = er.forward()
    # the forward run of every member against the oracle (its OpenMP build,
    # members spread over the host threads), bit for bit
    import os
    from oracle import oracle as O
    O.build()
    print("[c4] all 64 members' forward run against the oracle", flush=True)
    with O.libm("omp"):
        qos, _ = O.run_patches([O.params_from(m.patch.params) for m in members],
                               [m.patch.q for m in members], [m.patch.aux for m in members],
                               [m.patch.dx for m in members], [m.patch.dy for m in members],
                               nsteps, 1e-3, 0.9, 1e9, nthreads=min(16, os.cpu_count() or 1))
    I = (slice(None), slice(2, -2), slice(2, -2))
    for k in range(len(members)):
        assert np.array_equal(S.to_host_q(qf[k])[I], qos[k][I]), k
    del qos
    gs = [torch.zeros_like(s) for s in series]
    for k in range(len(members)):
        gs[k][:, :, 0] = 2.0 * (series[k][:, :, 0] - tg[k])
    lam0, grad = er.backward(gs, None)
    torch.cuda.synchronize()
    print(f"[c4] peak HBM {torch.cuda.max_memory_allocated() / 1e9:.1f} GB", flush=True)
    for k in (0, len(members) - 1):
        p = members[k].patch
        spec = AD.RolloutSpec(aux=p.aux, params=p.params, dx=p.dx, dy=p.dy, nsteps=nsteps,
                              gauges_xy=gxy[k], parameters=("phi", "kref"), checkpoint_every=6)
        r = AD.Rollout(spec)
        s1, q1 = r.forward(S.to_device_q(p.q), r.record())
        assert torch.equal(s1, series[k]) and torch.equal(q1, qf[k]), k
        g1 = torch.zeros_like(s1)
        g1[:, :, 0] = 2.0 * (s1[:, :, 0] - tg[k])
        l1, gr1 = r.backward(g1, None)
        torch.cuda.synchronize()
        assert torch.equal(l1, lam0[k]), k
        assert torch.equal(gr1, grad[k]), k
        del r
    # transpose: member 0's reverse gradient against its tangent-linear run
    p = members[0].patch
    rng = np.random.default_rng(5)
    dq_h = np.zeros_like(p.q)
    wet = p.q[0] > 0.0
    dq_h[0][wet] = rng.normal(size=wet.sum()) * 0.01
    targets = tg[0].cpu().numpy()
    loss_t, grads_t, _ = T.gauge_loss_gradient(
        p.q, p.aux, p.params, p.dx, p.dy, 0.0, 0.0, gxy[0], targets,
        {"phi": (T.param_direction(phi=1.0), None), "h": (None, dq_h)})
    lam_h = S.to_host_q(lam0[0])
    assert float(grad[0, 0]) == pytest.approx(grads_t["phi"], rel=1e-9)
    assert float((lam_h * dq_h).sum()) == pytest.approx(grads_t["h"], rel=1e-9)
    assert abs(grads_t["h"]) > 0.0 and abs(grads_t["phi"]) > 0.0
