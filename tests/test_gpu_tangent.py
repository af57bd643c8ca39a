"""GPU: the tangent-linear step (dclaw_stepgrid_tl_dev) against the host
emulator of the same body, bit for bit in values and tangents; its values
against the step; and a multi-step tangent run against central differences
of the GPU step on the same time grid."""
import ctypes

import numpy as np
import pytest

from dclaw_amd.params import CParams, to_c
from helpers import advance_oracle, slope_patch
from test_emu_tangent import _tl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dclaw_amd import _lib
    _lib.lib()
    return torch


@pytest.mark.parametrize("bc,seed", [((1, 1, 1, 1), 0), ((3, 3, 3, 3), 1), (None, 2)])
def test_tangent_step_bitwise_emulator(dev, oracle_mod, emu_lib, bc, seed):
    from dclaw_amd import _lib
    from dclaw_amd import step as S
    from dclaw_amd import tangent as T
    torch = dev
    pt = slope_patch(130, 70, slope=0.9, phi_deg=22.0, seed=11 + seed)
    q, aux, dt = advance_oracle(oracle_mod, pt, 8, bc=bc or (1, 1, 1, 1))
    q, aux = np.asfortranarray(q), np.asfortranarray(aux)
    rng = np.random.default_rng(seed)
    dq = np.asfortranarray(rng.normal(size=q.shape) * np.abs(q))
    dpar = T.param_direction(phi=1.0, kref=1e-10)
    print(f"[tangent] 130x70 bc {bc} seed {seed}", flush=True)
    L = _lib.lib()
    prev = L.dclaw_set_seg_rows(8)
    try:
        qe, dqe, cfle = _tl(emu_lib, to_c(pt.params), dpar, q, dq, aux, pt.dx, pt.dy, dt, 8, bc)
        qi, dqi = S.to_device_q(q), S.to_device_q(dq)
        qo, dqo = torch.zeros_like(qi), torch.zeros_like(qi)
        bed = S.bed_tensor(aux)
        cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
        dtd = torch.tensor([dt], dtype=torch.float64, device="cuda")
        T.stepgrid_tl(qi, dqi, qo, dqo, bed, 1, pt.mx, pt.my, pt.dx, pt.dy, bc, dtd,
                      S.params_tensor([pt.params]), T._record_tensor(dpar), cfl)
        # the step itself on the same input
        qf = torch.zeros_like(qi)
        cf = torch.zeros(1, dtype=torch.float64, device="cuda")
        if bc is None:
            S.stepgrid(qi, qf, bed, 1, pt.mx, pt.my, pt.dx, pt.dy, dtd, S.params_tensor([pt.params]),
                       cf)
        else:
            S.stepgrid_bc(qi, qf, bed, 1, pt.mx, pt.my, pt.dx, pt.dy, bc, dtd,
                          S.params_tensor([pt.params]), cf)
        torch.cuda.synchronize()
    finally:
        L.dclaw_set_seg_rows(prev)
    I = (slice(None), slice(2, -2), slice(2, -2))
    a, b = S.to_host_q(qo), S.to_host_q(dqo)
    assert np.array_equal(a[I], qe[I]) and cfl.item() == cfle
    assert np.array_equal(b[I], dqe[I]), f"{(b[I] != dqe[I]).sum()} tangents differ"
    assert np.array_equal(a, S.to_host_q(qf)) and cfl.item() == cf.item()
    assert np.abs(b[I]).max() > 0.0


def test_tangent_run_matches_finite_differences(dev):
    """30 steps of the runner (the values' own dt rule, recorded) on a
    fractal-bed patch with sources, d/dkref and d/dmu of the final state
    against central differences of the forward step on the same dt sequence."""
    from dclaw_amd import step as S
    from dclaw_amd import tangent as T
    from dclaw_amd import workloads as W
    torch = dev
    pt = W.fractal_sources(300, 200, dx=2.0, seed=6, nsrc=6)
    nsteps = 30
    base = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy)
    base.run(nsteps, 1e-3, 0.9, 1e9)
    for name in ("kref", "mu"):
        r = T.TangentRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy,
                            dparams=T.param_direction(**{name: 1.0}))
        r.run(nsteps, 1e-3, 0.9, 1e9, record_dt=True)
        qv, dq = r.q(), r.dq()
        assert np.array_equal(qv, base.q())          # values = the step's trajectory
        c0 = getattr(to_c(pt.params), name)
        eps = 1e-6 * c0

        def fwd(sign):
            p = to_c(pt.params)
            setattr(p, name, c0 + sign * eps)
            rr = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy)
            rr.params = torch.frombuffer(bytearray(bytes(p)), dtype=torch.uint8).to("cuda")
            for d in r.dts:
                rr.set_dt(d)
                rr.cfl.zero_()
                rr.step()
            return rr.q()
        fd = (fwd(+1) - fwd(-1)) / (2 * eps)
        I = (slice(None), slice(2, -2), slice(2, -2))
        for m in range(5):
            scale = np.abs(fd[I][m]).max()
            assert scale > 0.0
            floor = 8.0 * np.abs(qv[I][m]).max() * np.finfo(float).eps / eps
            err = np.abs(dq[I][m] - fd[I][m])
            nbad = int((err > 1e-3 * scale + floor).sum())
            assert nbad <= 0.001 * err.size, f"d/d{name} component {m}: {nbad} cells"


def test_gauge_loss_gradient_matches_finite_differences(dev):
    """Gradient of a gauge misfit (sum over steps and gauges of h^2) with
    respect to phi, kref (parameters) and m0 (initial-state direction: hm =
    m0 h) by tangent-linear runs, against central differences of the forward
    runs on the same time grid."""
    from dclaw_amd import step as S
    from dclaw_amd import tangent as T
    from dclaw_amd import workloads as W
    from dclaw_amd.params import to_c
    torch = dev
    pt, _ = W.flume_release(mx=400, my=12, dx=0.05)
    gxy = np.array([[0.9, 0.3], [1.3, 0.3], [1.7, 0.25], [2.2, 0.35]])
    nsteps = 120
    targets = np.zeros((nsteps, len(gxy)))
    dq_m0 = np.zeros_like(pt.q)
    dq_m0[3] = pt.q[0]
    dirs = {"phi": (T.param_direction(phi=1.0), None), "kref": (T.param_direction(kref=1.0), None),
            "m0": (None, dq_m0)}
    loss, grads, dts = T.gauge_loss_gradient(pt.q, pt.aux, pt.params, pt.dx, pt.dy, 0.0, 0.0, gxy,
                                             targets, dirs)
    assert loss > 0.0

    def fwd_loss(c=None, q0=None):
        r = S.PatchRunner(pt.q if q0 is None else q0, pt.aux, pt.params, pt.dx, pt.dy)
        if c is not None:
            r.params = torch.frombuffer(bytearray(bytes(c)), dtype=torch.uint8).to("cuda")
        xy = torch.tensor(gxy, device="cuda")
        out = torch.zeros((len(gxy), 8), dtype=torch.float64, device="cuda")
        tot = 0.0
        for d in dts:
            r.set_dt(float(d))
            r.cfl.zero_()
            r.step()
            S.gauges(r.qa, r.bed, r.maux, r.mx, r.my, 0.0, 0.0, r.dx, r.dy,
                     pt.params.dry_tolerance, xy, out)
            tot += float((out[:, 0] ** 2).sum().item())
        return tot
    assert fwd_loss() == pytest.approx(loss, rel=1e-14)
    for name in ("phi", "kref"):
        c0 = getattr(to_c(pt.params), name)
        eps = 1e-6 * c0
        cp, cm = to_c(pt.params), to_c(pt.params)
        setattr(cp, name, c0 + eps)
        setattr(cm, name, c0 - eps)
        fd = (fwd_loss(c=cp) - fwd_loss(c=cm)) / (2 * eps)
        assert grads[name] == pytest.approx(fd, rel=1e-4), name
    eps = 1e-6
    fd = (fwd_loss(q0=pt.q + eps * dq_m0) - fwd_loss(q0=pt.q - eps * dq_m0)) / (2 * eps)
    assert grads["m0"] == pytest.approx(fd, rel=1e-4)


def test_flume_tangent_is_the_derivative_of_the_discrete_map(dev):
    """On the 2000 x 40 flume release the tangent after 150 steps equals
    central differences taken at a tiny step (1e-8 relative in phi): the
    map's linear regime is that narrow there (larger steps cross the
    limiter's and the friction law's kinks), and the tangent is its
    derivative."""
    from dclaw_amd import _lib
    from dclaw_amd import step as S
    from dclaw_amd import tangent as T
    from dclaw_amd import workloads as W
    torch = dev
    L = _lib.lib()
    prev = L.dclaw_set_seg_rows(4)
    try:
        pt, _ = W.flume_release()
        r = T.TangentRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy,
                            dparams=T.param_direction(phi=1.0))
        r.run(150, 1e-3, 0.9, 1e9, record_dt=True)
        dq = r.dq()
        c0 = to_c(pt.params).phi
        eps = 1e-8 * c0

        def fwd(sign):
            p = to_c(pt.params)
            p.phi = c0 + sign * eps
            rr = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy)
            rr.params = torch.frombuffer(bytearray(bytes(p)), dtype=torch.uint8).to("cuda")
            for d in r.dts:
                rr.set_dt(d)
                rr.cfl.zero_()
                rr.step()
            return rr.q()
        fd = (fwd(1) - fwd(-1)) / (2 * eps)
    finally:
        L.dclaw_set_seg_rows(prev)
    I = (slice(None), slice(2, -2), slice(2, -2))
    for m in (0, 1, 3, 4):
        scale = np.abs(dq[I][m]).max()
        assert scale > 0.0
        assert np.abs(fd[I][m] - dq[I][m]).max() <= 1e-4 * scale, m


def test_batched_tangent_equals_per_member_launches(dev):
    """Five members of uneven sizes with different directions: the batched
    tangent-linear launch equals the per-member launches bit for bit (values
    and tangents), and its values equal the forward ensemble's."""
    from dclaw_amd import step as S
    from dclaw_amd import tangent as T
    from dclaw_amd import workloads as W
    torch = dev
    members = W.ensemble(5, 60, 200, seed=4)
    dirs = [T.param_direction(phi=1.0), T.param_direction(kref=1e-10), T.param_direction(mu=1.0),
            T.param_direction(phi=0.5, delta=1.0), T.param_direction()]
    a = T.TangentEnsemble(members, directions=dirs, seg_rows=8)
    b = T.TangentEnsemble(members, directions=dirs, seg_rows=8)
    f = S.Ensemble(members, seg_rows=8, split_src2=False)
    a.run(12, 1e-3, 0.9, 1e9)
    b.run(12, 1e-3, 0.9, 1e9, plain=True)
    f.run(12, 1e-3, 0.9, 1e9)
    torch.cuda.synchronize()
    for k in range(5):
        assert np.array_equal(a.q(k), b.q(k)) and np.array_equal(a.dq(k), b.dq(k)), k
        assert np.array_equal(a.q(k), f.q(k)), k
    assert np.abs(a.dq(0)).max() > 0.0 and not a.dq(4).any()
