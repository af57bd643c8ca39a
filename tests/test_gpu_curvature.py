"""GPU: src2's curvature term (each wet cell's gravity includes the curvature
accelerations of the cells before it in src2's loop order) through
dclaw_stepgrid_curvature_dev (PatchRunner), dclaw_src2_curvature_dev and the
host drop-in src2_, against the oracle bit for bit; the other stepping entry
points mark wet cells NaN when curvature is on."""
import ctypes

import numpy as np
import pytest

from helpers import advance_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dclaw_amd import _lib
    _lib.lib()
    return torch


def _valley(mx=96, my=64, seed=4):
    """A mound released into a curved valley (bed quadratic across, tilted
    along, rippled), so the curvature term is active on most wet cells."""
    from dclaw_amd import workloads as W
    rng = np.random.default_rng(seed)
    p = W.config0_params()
    p.phi_deg = 25.0
    p.curvature = 1
    x = np.arange(mx) + 0.5
    y = np.arange(my) + 0.5
    X, Y = np.meshgrid(x, y, indexing="ij")
    b = 0.6 * (mx - X) + 0.01 * (Y - my / 2) ** 2 + 0.3 * np.sin(0.15 * X) \
        + 0.02 * rng.standard_normal((mx, my))
    h = 3.0 * np.exp(-((X - 0.3 * mx) ** 2 + (Y - 0.5 * my) ** 2) / (2 * 8.0 ** 2))
    return W.make_patch(b, h, 0.62, 0.7, p, 1.0, 1.2)


def test_curvature_runner_equals_oracle(dev, oracle_mod):
    from dclaw_amd import step as S
    from helpers import mask_equal
    O = oracle_mod
    torch = dev
    pt = _valley()
    n = 40
    print("[curvature] 96x64 valley, 40 steps", flush=True)
    qo, _, dto, _, _ = O.run(O.params_from(pt.params), pt.q, pt.aux, pt.dx, pt.dy, n, 1e-3, 0.9,
                             1e9)
    r = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy)
    assert r.curv_work is not None and r.maux == pt.aux.shape[0]
    r.run(n, 1e-3, 0.9, 1e9)
    torch.cuda.synchronize()
    I = (slice(None), slice(2, -2), slice(2, -2))
    qg = r.q()
    assert mask_equal(qg[I], qo[I], pt.params.dry_tolerance)
    assert np.array_equal(qg[I], qo[I])
    assert r.dt.item() == dto
    # the term matters here: the same run without it differs
    pt.params.curvature = 0
    q0, _, _, _, _ = O.run(O.params_from(pt.params), pt.q, pt.aux, pt.dx, pt.dy, n, 1e-3, 0.9,
                           1e9)
    assert not np.array_equal(q0[I], qo[I])


def test_curvature_src2_entry_points(dev, oracle_mod):
    """dclaw_src2_curvature_dev and src2_ (Fortran convention, host memory)
    equal the oracle's src2; the plain src2 entry point marks wet pb NaN."""
    from dclaw_amd import _lib
    from dclaw_amd import step as S
    from dclaw_amd.params import to_c
    O = oracle_mod
    torch = dev
    L = _lib.lib()
    pt = _valley(70, 50, seed=6)
    pt.params.curvature = 0
    q, aux, dt = advance_oracle(O, pt, 8)
    pt.params.curvature = 1
    P = O.params_from(pt.params)
    want = O.src2(P, q, aux, dt, dx=pt.dx, dy=pt.dy)
    mx, my = pt.mx, pt.my
    qd = S.to_device_q(q)
    ad = S.aux_tensor(aux)
    work = torch.zeros(mx * my, dtype=torch.float64, device="cuda")
    dtd = torch.tensor([dt], dtype=torch.float64, device="cuda")
    prm = S.params_tensor([pt.params])
    _lib.check(L.dclaw_src2_curvature_dev(S._ptr(qd), S._ptr(ad), aux.shape[0], mx, my, pt.dx,
                                          pt.dy, S._ptr(dtd), S._ptr(prm), S._ptr(work),
                                          work.numel(), S._stream()))
    torch.cuda.synchronize()
    assert np.array_equal(S.to_host_q(qd), want)
    # Fortran-convention drop-in
    _lib.check(L.dclaw_set_params(ctypes.byref(to_c(pt.params))))
    dp = ctypes.POINTER(ctypes.c_double)
    ci = lambda v: ctypes.byref(ctypes.c_int(v))
    cd = lambda v: ctypes.byref(ctypes.c_double(v))
    qh = np.array(q, order="F", copy=True)
    ah = np.array(aux, order="F", copy=True)
    L.src2_(ci(7), ci(2), ci(mx), ci(my), cd(0.0), cd(0.0), cd(pt.dx), cd(pt.dy),
            qh.ctypes.data_as(dp), ci(aux.shape[0]), ah.ctypes.data_as(dp), cd(0.0), cd(dt))
    assert np.array_equal(qh, want)
    # the plain per-stage src2 cannot form the running gravity: NaN, loudly
    qd = S.to_device_q(q)
    S.src2(qd, ad, mx, my, dtd, prm)
    torch.cuda.synchronize()
    out = S.to_host_q(qd)
    wet = want[0][2:-2, 2:-2] > pt.params.dry_tolerance
    assert np.isnan(out[4][2:-2, 2:-2][wet]).all()
