"""GPU: the adjoint step (dclaw_adjoint_step_dev) against the host emulator
of the same stages; the transpose identity against the GPU tangent-linear
step; and gradients of multi-step gauge losses by the checkpointed reverse
sweep (and through torch.autograd) against the tangent-linear gradients
(which test_gpu_tangent.py pins to finite differences) and directly against
central differences of the forward runs."""
import numpy as np
import pytest

from dclaw_amd.params import CParams, to_c
from helpers import advance_oracle, slope_patch
from test_emu_adjoint import _adj, _scaled

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dclaw_amd import _lib
    _lib.lib()
    return torch


def _gpu_adjoint(torch, pt, q, aux, dt, bc, lam, dPs):
    from dclaw_amd import adjoint as AD
    from dclaw_amd import step as S
    from dclaw_amd import tangent as T
    mx, my = pt.mx, pt.my
    qi = S.to_device_q(q)
    li = S.to_device_q(lam)
    lo = torch.zeros_like(qi)
    qo = torch.zeros_like(qi)
    bed = S.bed_tensor(aux)
    work = AD.new_workspace(mx, my)
    dtd = torch.tensor([dt], dtype=torch.float64, device="cuda")
    dpar = torch.cat([T._record_tensor(d) for d in dPs]) if dPs else None
    grad = torch.zeros(max(1, len(dPs)), dtype=torch.float64, device="cuda")
    AD.adjoint_step(qi, bed, 1, mx, my, pt.dx, pt.dy, bc, dtd, S.params_tensor([pt.params]),
                    dpar, len(dPs), li, lo, qo, grad if dPs else None, work)
    torch.cuda.synchronize()
    return S.to_host_q(qo), S.to_host_q(lo), grad.cpu().numpy()[:len(dPs)]


@pytest.mark.parametrize("bc,seed", [((1, 1, 1, 1), 0), ((3, 3, 3, 3), 1), (None, 2),
                                     ((1, 3, 3, 1), 3)])
def test_adjoint_step_equals_emulator(dev, oracle_mod, emu_lib, bc, seed):
    torch = dev
    pt = slope_patch(130, 70, slope=0.9, phi_deg=22.0, seed=11 + seed)
    q, aux, dt = advance_oracle(oracle_mod, pt, 8, bc=bc or (1, 1, 1, 1))
    q, aux = np.asfortranarray(q), np.asfortranarray(aux)
    rng = np.random.default_rng(seed)
    lam = np.asfortranarray(rng.normal(size=q.shape))
    dPs = [CParams(phi=1.0), CParams(kref=1e-10)]
    print(f"[adjoint] 130x70 bc {bc} seed {seed}", flush=True)
    qe, le, ge = _adj(emu_lib, to_c(pt.params), dPs, q, aux, pt.dx, pt.dy, dt, bc, lam)
    qg, lg, gg = _gpu_adjoint(torch, pt, q, aux, dt, bc, lam, dPs)
    assert np.array_equal(qg, qe), "record pass differs"
    for m in range(7):
        sc = np.abs(le[m]).max()
        assert np.abs(lg[m] - le[m]).max() <= 1e-13 * sc, m
    assert np.allclose(gg, ge, rtol=1e-12, atol=0.0)
    assert np.abs(lg).max() > 0.0


def test_adjoint_is_transpose_of_gpu_tangent(dev):
    """<lambda, TL(dq, dP)> = <J^T lambda, dq> + g . a on a fractal-bed patch
    with sources (GPU tangent-linear step against GPU adjoint step)."""
    from dclaw_amd import step as S
    from dclaw_amd import tangent as T
    from dclaw_amd import workloads as W
    torch = dev
    pt = W.fractal_sources(300, 200, dx=2.0, seed=6, nsrc=6)
    r = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy)
    r.run(20, 1e-3, 0.9, 1e9)
    q = r.q()
    dt = float(r.dt.item())
    rng = np.random.default_rng(4)
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


def _flume_case():
    from dclaw_amd import workloads as W
    pt, _ = W.flume_release(mx=400, my=12, dx=0.05)
    gxy = np.array([[0.9, 0.3], [1.3, 0.3], [1.7, 0.25], [2.2, 0.35]])
    nsteps = 120
    rng = np.random.default_rng(2)
    targets = 0.1 * rng.random((nsteps, len(gxy)))
    return pt, gxy, nsteps, targets


def test_reverse_gradient_equals_tangent_gradients(dev):
    """Gauge misfit over 120 steps of a flume release: the one reverse sweep
    gives d/dphi and d/dkref equal to the tangent-linear runs', and its
    initial-state gradient dotted with the m0 direction (hm = m0 h) and with
    a random depth perturbation equals the tangent-linear derivatives along
    them; any checkpoint interval and sub-checkpoint interval give the same
    result bit for bit."""
    from dclaw_amd import adjoint as AD
    from dclaw_amd import tangent as T
    pt, gxy, nsteps, targets = _flume_case()
    dq_m0 = np.zeros_like(pt.q)
    dq_m0[3] = pt.q[0]
    rng = np.random.default_rng(5)
    dq_h = np.zeros_like(pt.q)
    wet = pt.q[0] > 0.0
    dq_h[0][wet] = rng.normal(size=wet.sum()) * 0.01
    dirs = {"phi": (T.param_direction(phi=1.0), None), "kref": (T.param_direction(kref=1.0), None),
            "m0": (None, dq_m0), "h": (None, dq_h)}
    loss_t, grads_t, dts_t = T.gauge_loss_gradient(pt.q, pt.aux, pt.params, pt.dx, pt.dy, 0.0, 0.0,
                                                   gxy, targets, dirs)
    loss_a, grads_a, lam0, dts_a = AD.gauge_loss_gradient_adjoint(
        pt.q, pt.aux, pt.params, pt.dx, pt.dy, 0.0, 0.0, gxy, targets, parameters=("phi", "kref"),
        checkpoint_every=7)
    assert np.array_equal(dts_a, dts_t)
    assert loss_a == pytest.approx(loss_t, rel=1e-14)
    for n in ("phi", "kref"):
        assert grads_a[n] == pytest.approx(grads_t[n], rel=1e-9), n
    assert float((lam0 * dq_m0).sum()) == pytest.approx(grads_t["m0"], rel=1e-9)
    assert float((lam0 * dq_h).sum()) == pytest.approx(grads_t["h"], rel=1e-9)
    for every, sub in [(1000, 0), (1000, 1), (7, 3), (40, 40)]:
        _, grads_b, lam0_b, _ = AD.gauge_loss_gradient_adjoint(
            pt.q, pt.aux, pt.params, pt.dx, pt.dy, 0.0, 0.0, gxy, targets,
            parameters=("phi", "kref"), checkpoint_every=every, sub_every=sub)
        assert np.array_equal(lam0_b, lam0) and grads_b == grads_a, (every, sub)


def test_reverse_gradient_matches_finite_differences(dev):
    """The reverse sweep's gradient of a gauge misfit over 120 steps of a
    flume release (checkpoints every 10 steps) against central differences of
    the forward runs on the same time grid: d/dphi and d/dkref at a 1e-6
    relative step, and the initial-state gradient along the m0 direction
    (hm = m0 h) and along a uniform scaling of the initial depth field."""
    from dclaw_amd import adjoint as AD
    from dclaw_amd import step as S
    torch = dev
    pt, gxy, nsteps, targets = _flume_case()
    loss, grads, lam0, dts = AD.gauge_loss_gradient_adjoint(
        pt.q, pt.aux, pt.params, pt.dx, pt.dy, 0.0, 0.0, gxy, targets, parameters=("phi", "kref"),
        checkpoint_every=10)

    def fwd_loss(c=None, q0=None):
        r = S.PatchRunner(pt.q if q0 is None else q0, pt.aux, pt.params, pt.dx, pt.dy)
        if c is not None:
            r.params = torch.frombuffer(bytearray(bytes(c)), dtype=torch.uint8).to("cuda")
        xy = torch.tensor(gxy, device="cuda")
        out = torch.zeros((len(gxy), 8), dtype=torch.float64, device="cuda")
        tot = 0.0
        for n, d in enumerate(dts):
            r.set_dt(float(d))
            r.cfl.zero_()
            r.step()
            S.gauges(r.qa, r.bed, r.maux, r.mx, r.my, 0.0, 0.0, r.dx, r.dy,
                     pt.params.dry_tolerance, xy, out)
            res = out[:, 0].cpu().numpy() - targets[n]
            tot += float((res * res).sum())
        return tot

    assert fwd_loss() == pytest.approx(loss, rel=1e-13)
    for name in ("phi", "kref"):
        c0 = getattr(to_c(pt.params), name)
        eps = 1e-6 * c0
        cp, cm = to_c(pt.params), to_c(pt.params)
        setattr(cp, name, c0 + eps)
        setattr(cm, name, c0 - eps)
        fd = (fwd_loss(c=cp) - fwd_loss(c=cm)) / (2 * eps)
        print(f"[adjoint fd] {name}: reverse {grads[name]:.10e} central {fd:.10e}", flush=True)
        assert grads[name] == pytest.approx(fd, rel=1e-4), name
    dq_m0 = np.zeros_like(pt.q)
    dq_m0[3] = pt.q[0]
    eps = 1e-6
    fd = (fwd_loss(q0=pt.q + eps * dq_m0) - fwd_loss(q0=pt.q - eps * dq_m0)) / (2 * eps)
    g_m0 = float((lam0 * dq_m0).sum())
    print(f"[adjoint fd] m0: reverse {g_m0:.10e} central {fd:.10e}", flush=True)
    assert g_m0 == pytest.approx(fd, rel=1e-4)
    # the initial depth field, scaled uniformly (a smooth direction: a random
    # cell-by-cell perturbation turns neighbouring slopes over, so central
    # differences at this step size cross the limiter's kinks)
    dq_h = np.zeros_like(pt.q)
    dq_h[0] = pt.q[0]
    fd = (fwd_loss(q0=pt.q + eps * dq_h) - fwd_loss(q0=pt.q - eps * dq_h)) / (2 * eps)
    g_h = float((lam0 * dq_h).sum())
    print(f"[adjoint fd] h0: reverse {g_h:.10e} central {fd:.10e}", flush=True)
    assert g_h == pytest.approx(fd, rel=1e-4)


def test_autograd_rollout(dev):
    """torch.autograd through ``rollout``: q0 built from a depth field h0 and
    a solid fraction m0 by torch operations, a loss on the gauge records and
    on the final state (a runout-like first moment of the depth); the
    gradients with respect to h0, m0 and (phi, kref) equal the
    tangent-linear derivatives of the same loss."""
    from dclaw_amd import adjoint as AD
    from dclaw_amd import step as S
    from dclaw_amd import tangent as T
    torch = dev
    pt, gxy, nsteps, targets = _flume_case()
    nsteps = 60
    targets = targets[:nsteps]
    spec = AD.RolloutSpec(aux=pt.aux, params=pt.params, dx=pt.dx, dy=pt.dy, nsteps=nsteps,
                          gauges_xy=gxy, parameters=("phi", "kref"), checkpoint_every=16)
    base = S.to_device_q(pt.q)
    h0 = base[:, :, 0].clone().requires_grad_(True)
    m0 = torch.tensor(0.6, dtype=torch.float64, device="cuda", requires_grad=True)
    c = to_c(pt.params)
    theta = torch.tensor([c.phi, c.kref], dtype=torch.float64, requires_grad=True)
    ratio = base[:, :, 4] / torch.where(base[:, :, 0] > 0, base[:, :, 0], 1.0)  # pb / h, kept

    def build(h, m):
        cols = [h, base[:, :, 1], base[:, :, 2], m * h, ratio * h, base[:, :, 5], base[:, :, 6]]
        return torch.stack(cols, dim=2)
    q0 = build(h0, m0)
    series, qf = AD.rollout(q0, theta, spec)
    xs = (torch.arange(pt.mx + 4, dtype=torch.float64, device="cuda") - 1.5) * pt.dx
    tgt = torch.tensor(targets, device="cuda")
    loss = ((series[:, :, 0] - tgt) ** 2).sum() + 0.1 * (qf[:, :, 0] * xs[None, :]).sum()
    gh, gm, gth = torch.autograd.grad(loss, [h0, m0, theta])
    # tangent-linear derivatives of the same loss along (dh, dm0, dtheta)
    rng = np.random.default_rng(8)
    dh = torch.zeros_like(h0)
    wet = base[:, :, 0] > 0
    dh[wet] = torch.tensor(rng.normal(size=int(wet.sum().item())) * 0.01, device="cuda")

    q0_host = S.to_host_q(q0.detach())

    def tl_loss(dq0_dev, dparams):
        r = T.TangentRunner(q0_host, pt.aux, pt.params, pt.dx, pt.dy, dparams=dparams)
        r.dqa.copy_(dq0_dev)
        xy = torch.tensor(gxy, device="cuda")
        out = torch.zeros((len(gxy), 8), dtype=torch.float64, device="cuda")
        dout = torch.zeros_like(out)
        r.set_dt(1e-3)
        r.cfl.zero_()
        acc = torch.zeros(1, dtype=torch.float64, device="cuda")
        for n in range(nsteps):
            r.step()
            S.dt_update(r.dt, r.cfl, 0.9, 1e9)
            T.gauges_tl(r.qa, r.dqa, r.aux, r.maux, r.mx, r.my, 0.0, 0.0, r.dx, r.dy,
                        pt.params.dry_tolerance, xy, out, dout)
            acc += (2.0 * (out[:, 0] - tgt[n]) * dout[:, 0]).sum()
        acc += 0.1 * (r.dqa[:, :, 0] * xs[None, :]).sum()
        return float(acc.item())
    with torch.no_grad():
        dq_h = build(dh, m0.detach()) - build(torch.zeros_like(dh), m0.detach())
        dq_m = build(h0.detach(), torch.ones_like(m0)) - build(h0.detach(), torch.zeros_like(m0))
    assert float((gh * dh).sum()) == pytest.approx(tl_loss(dq_h, T.param_direction()), rel=1e-9)
    assert float(gm) == pytest.approx(tl_loss(dq_m, T.param_direction()), rel=1e-9)
    z = torch.zeros_like(base)
    assert float(gth[0]) == pytest.approx(tl_loss(z, T.param_direction(phi=1.0)), rel=1e-9)
    assert float(gth[1]) == pytest.approx(tl_loss(z, T.param_direction(kref=1.0)), rel=1e-9)


def test_sparse_tiles_and_workspace_reuse(dev, emu_lib):
    """Small wet blobs in a 200 x 150 dry patch with random bdif (the tile
    passes cover a few tiles; elsewhere only the dry map's bdif adjoint, with
    the border band's dry ghost images gathered by the dense pass): one
    workspace serves four calls (blob A, blob B elsewhere, A again, a blob
    across the wall near a corner), each equal to the emulator's dense
    stages, the third bit for bit the first."""
    from dclaw_amd import adjoint as AD
    from dclaw_amd import step as S
    from dclaw_amd import tangent as T
    torch = dev
    pt = slope_patch(200, 150, slope=0.9, phi_deg=22.0, seed=3)
    rng = np.random.default_rng(9)
    X, Y = np.meshgrid(np.arange(-1, 203) - 0.5, np.arange(-1, 153) - 0.5, indexing="ij")

    def blob(cx, cy, r):
        q = np.array(pt.q, order="F", copy=True)
        out = (X - cx) ** 2 + (Y - cy) ** 2 > r * r
        for m in range(6):
            q[m][out] = 0.0
        q[0][~out] = np.maximum(q[0][~out], 0.5)
        q[1][~out] = 0.3 * q[0][~out]
        q[6] = 0.01 * rng.standard_normal(q[6].shape)
        return q
    states = [blob(50, 40, 10), blob(150, 110, 14)]
    states.append(states[0])
    # a blob over the walled x-high edge near a corner: border tiles beside it
    # run the passes, the rest of the band keeps the dense pass's dry gather
    states.append(blob(197, 4, 9))
    aux = np.asfortranarray(pt.aux)
    bc = (1, 3, 1, 1)
    dPs = [CParams(phi=1.0), CParams(kref=1e-10)]
    work = AD.new_workspace(200, 150)
    bed = S.bed_tensor(aux)
    par = S.params_tensor([pt.params])
    dpar = torch.cat([T._record_tensor(d) for d in dPs])
    dtd = torch.tensor([0.02], dtype=torch.float64, device="cuda")
    lam = np.asfortranarray(rng.normal(size=pt.q.shape))
    got = []
    for k, q in enumerate(states):
        print(f"[adjoint sparse] call {k}", flush=True)
        qi, li = S.to_device_q(q), S.to_device_q(lam)
        lo, qo = torch.zeros_like(qi), torch.zeros_like(qi)
        grad = torch.zeros(2, dtype=torch.float64, device="cuda")
        AD.adjoint_step(qi, bed, 1, 200, 150, pt.dx, pt.dy, bc, dtd, par, dpar, 2, li, lo, qo, grad,
                        work)
        torch.cuda.synchronize()
        qe, le, ge = _adj(emu_lib, to_c(pt.params), dPs, q, aux, pt.dx, pt.dy, 0.02, bc, lam)
        qg, lg, gg = S.to_host_q(qo), S.to_host_q(lo), grad.cpu().numpy()
        assert np.array_equal(qg, qe), k
        for m in range(7):
            sc = np.abs(le[m]).max()
            assert np.abs(lg[m] - le[m]).max() <= 1e-13 * sc, (k, m)
        assert np.allclose(gg, ge, rtol=1e-12, atol=0.0), k
        got.append((lg, gg))
    assert np.array_equal(got[2][0], got[0][0]) and np.array_equal(got[2][1], got[0][1])
    # the dry map's adjoint far from the blobs: only bdif, where bdif > 0
    far = (slice(None), slice(100, 120), slice(10, 30))
    assert not got[0][0][:6][far[1:]].any()
    assert np.array_equal(got[0][0][6][far[1:]], np.where(states[0][6][far[1:]] > 0,
                                                          lam[6][far[1:]], 0.0))


@pytest.mark.parametrize("mixed", [False, True])
def test_batched_adjoint_equals_per_patch(dev, mixed):
    """Six members of uneven sizes, each with its own parameters and dt: the
    batched adjoint step (one launch per pass) equals the per-member
    single-patch adjoint steps bit for bit (lambda_n and the phi, kref
    gradients), on two successive states with the same workspaces.  mixed:
    two members with Manning friction, whose src2 reverse takes the seeded
    products while the others take the written-out one."""
    from dclaw_amd import adjoint as AD
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    torch = dev
    members = W.ensemble(6, 60, 200, seed=4)
    if mixed:
        for k in (1, 4):
            members[k].patch.params.friction_forcing = 1
            members[k].patch.params.manning_coefficient = 0.06
    ens = S.Ensemble(members, seg_rows=16)
    ens.run(8, 1e-3, 0.9, 1e9)
    adj = AD.EnsembleAdjoint(ens, parameters=("phi", "kref"))
    g = torch.Generator(device="cuda").manual_seed(1)
    for call in range(2):
        print(f"[adjoint batched] mixed {mixed} call {call}", flush=True)
        q = [ens.current(k).clone() for k in range(ens.n)]
        lam = [torch.randn(x.shape, dtype=torch.float64, device="cuda", generator=g) for x in q]
        lb = [torch.zeros_like(x) for x in q]
        lp = [torch.zeros_like(x) for x in q]
        gb = adj.step(q, lam, lb).clone()
        gp = adj.step_plain(q, lam, lp).clone()
        torch.cuda.synchronize()
        for k in range(ens.n):
            assert torch.equal(lb[k], lp[k]), (call, k)
        assert torch.equal(gb, gp), call
        assert float(gb.abs().max()) > 0.0
        ens.run(5, ens.dt.clone(), 0.9, 1e9)


def test_ensemble_rollout_equals_member_rollouts(dev):
    """Four members of uneven sizes, each with its own gauges and targets:
    the batched reverse sweep (EnsembleRollout, checkpoints every 5 steps)
    gives each member's gauge-misfit gradient (lambda_0 and d/dphi, d/dkref)
    bit for bit as that member's own Rollout."""
    from dclaw_amd import adjoint as AD
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    torch = dev
    members = W.ensemble(4, 60, 140, seed=9)
    nsteps = 12
    rng = np.random.default_rng(3)
    gxy, tg = [], []
    for m in members:
        p = m.patch
        xy = np.column_stack([rng.uniform(2.0, (p.mx - 2) * p.dx, 3),
                              rng.uniform(2.0, (p.my - 2) * p.dy, 3)])
        gxy.append(xy)
        tg.append(torch.tensor(rng.uniform(0.0, 0.5, (nsteps, 3)), device="cuda"))
    er = AD.EnsembleRollout(members, nsteps, gxy, parameters=("phi", "kref"), checkpoint_every=5)
    series, qf = er.forward()
    gs = [torch.zeros_like(s) for s in series]
    for k in range(len(members)):
        gs[k][:, :, 0] = 2.0 * (series[k][:, :, 0] - tg[k])
    lam0, grad = er.backward(gs, None)
    torch.cuda.synchronize()
    for k, m in enumerate(members):
        p = m.patch
        spec = AD.RolloutSpec(aux=p.aux, params=p.params, dx=p.dx, dy=p.dy, nsteps=nsteps,
                              gauges_xy=gxy[k], parameters=("phi", "kref"), checkpoint_every=5)
        r = AD.Rollout(spec)
        s1, q1 = r.forward(S.to_device_q(p.q), r.record())
        assert torch.equal(s1, series[k]) and torch.equal(q1, qf[k]), k
        g1 = torch.zeros_like(s1)
        g1[:, :, 0] = 2.0 * (s1[:, :, 0] - tg[k])
        l1, gr1 = r.backward(g1, None)
        torch.cuda.synchronize()
        assert torch.equal(l1, lam0[k]), k
        assert torch.equal(gr1, grad[k]), k
    assert float(grad.abs().max()) > 0.0


def test_full_flume_reverse_gradient_matches_finite_differences(dev):
    """The 2000 x 40 flume release with four gauges on the centre line
    within reach of the flow (2-8 m; depth misfit against a seeded target),
    150 steps: the reverse sweep's
    dloss/dphi equals central differences of the loss over the same time grid
    at a 1e-8 relative step.  150 steps is the horizon over which the
    derivative of this run is validated; beyond it the derivative of the
    MC-limited second-order correction grows by orders of magnitude per few
    steps in the supercritical flow, and no finite difference resolves it."""
    from dclaw_amd import _lib
    from dclaw_amd import adjoint as AD
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    from dclaw_amd.params import to_c
    torch = dev
    L = _lib.lib()
    prev = L.dclaw_set_seg_rows(4)
    try:
        pt, _ = W.flume_release(seed=3)
        gxy = np.array([[2.0, 1.0], [4.0, 1.0], [6.0, 1.0], [8.0, 1.0]])
        nsteps = 150
        rng = np.random.default_rng(3)
        centres = rng.uniform(0.1 * nsteps, 0.9 * nsteps, len(gxy))
        n = np.arange(1, nsteps + 1)[:, None]
        targets = 0.3 * np.exp(-((n - centres[None, :]) / (0.1 * nsteps)) ** 2)
        print("[flume fd] reverse gradient over 150 steps", flush=True)
        loss, grads, _, dts = AD.gauge_loss_gradient_adjoint(
            pt.q, pt.aux, pt.params, pt.dx, pt.dy, 0.0, 0.0, gxy, targets,
            parameters=("phi",), checkpoint_every=50)
        c0 = to_c(pt.params).phi
        eps = 1e-8 * c0
        xy = torch.tensor(np.asarray(gxy, dtype=np.float64), device="cuda")
        tgt = torch.tensor(targets, device="cuda")

        def loss_at(phi):
            p = to_c(pt.params)
            p.phi = phi
            rr = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy)
            rr.params = torch.frombuffer(bytearray(bytes(p)), dtype=torch.uint8).to("cuda")
            rec = torch.zeros((nsteps, len(gxy), 8), dtype=torch.float64, device="cuda")
            for k, d in enumerate(dts):
                rr.set_dt(float(d))
                rr.cfl.zero_()
                rr.step()
                S.gauges(rr.qa, rr.bed, rr.maux, rr.mx, rr.my, 0.0, 0.0, rr.dx, rr.dy,
                         pt.params.dry_tolerance, xy, rec[k])
            res = rec[:, :, 0] - tgt
            return float((res * res).sum())
        assert loss_at(c0) == pytest.approx(loss, rel=1e-14)
        fd = (loss_at(c0 + eps) - loss_at(c0 - eps)) / (2 * eps)
    finally:
        L.dclaw_set_seg_rows(prev)
    assert abs(grads["phi"]) > 0.0
    assert grads["phi"] == pytest.approx(fd, rel=1e-4), (grads["phi"], fd)
