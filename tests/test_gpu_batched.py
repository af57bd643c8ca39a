"""GPU: batched ensembles (one launch per step over a work list), boundary
conditions on load, and the Fortran-convention host drop-in, against the
oracle and against the single-patch path."""
import ctypes

import numpy as np
import pytest

from helpers import TOL_LONG, TOL_STEP, mask_equal, rel_linf_wet, slope_patch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dclaw_amd import _lib
    _lib.lib()
    return torch


def _members(n=6, seed=5):
    from dclaw_amd import workloads as W
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        mx, my = int(rng.integers(10, 140)), int(rng.integers(10, 140))
        pt = slope_patch(mx, my, slope=float(rng.uniform(0.3, 0.9)),
                         phi_deg=float(rng.uniform(22, 35)), seed=k)
        pt.params.kref = float(10 ** rng.uniform(-11, -9))
        out.append(W.Member(pt, seed=k))
    return out


def test_batched_matches_oracle_per_member(dev, oracle_mod):
    from dclaw_amd import step as S
    O = oracle_mod
    members = _members()
    nsteps = 25
    print("[batched] sizes", [(m.patch.mx, m.patch.my) for m in members], flush=True)
    ens = S.Ensemble(members, seg_rows=16)
    ens.run(nsteps, 1e-3, 0.9, 1e9)
    dev.cuda.synchronize()
    for k, m in enumerate(members):
        pt = m.patch
        qo, _, dto, _, _ = O.run(O.params_from(pt.params), pt.q, pt.aux, pt.dx, pt.dy, nsteps,
                                 1e-3, 0.9, 1e9)
        qg = ens.q(k)
        I = (slice(None), slice(2, -2), slice(2, -2))
        assert mask_equal(qg[I], qo[I], pt.params.dry_tolerance), k
        assert rel_linf_wet(qg[I], qo[I], pt.params.dry_tolerance) <= TOL_LONG, k
        assert ens.dt[k].item() == pytest.approx(dto, rel=1e-12)


def test_batched_equals_plain_loop(dev):
    from dclaw_amd import step as S
    members = _members(seed=8)
    a = S.Ensemble(members, seg_rows=32)
    b = S.Ensemble(members, seg_rows=32)
    a.run(10, 1e-3, 0.9, 1e9)
    b.run(10, 1e-3, 0.9, 1e9, plain=True)
    dev.cuda.synchronize()
    for k in range(len(members)):
        assert np.array_equal(a.q(k)[:, 2:-2, 2:-2], b.q(k)[:, 2:-2, 2:-2]), k
    assert dev.equal(a.dt, b.dt)


def test_c2_ensemble_256(dev, oracle_mod):
    """The 256-member ensemble (sides log-uniform in [64, 512], per-member
    fractal bed, Gaussian source, phi, m0, kref): every member equals the
    oracle bit for bit after 25 batched steps (the oracle's OpenMP build,
    one member per thread at a time), dt included, and equals its own
    single-patch steps (the plain per-member loop)."""
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    O = oracle_mod
    members = W.ensemble(256, 64, 512, seed=1)
    nsteps = 25
    cells = [m.patch.mx * m.patch.my for m in members]
    print(f"[c2] 256 members, {sum(cells)} cells", flush=True)
    a = S.Ensemble(members, seg_rows=16)
    b = S.Ensemble(members, seg_rows=16)
    a.run(nsteps, 1e-3, 0.9, 1e9)
    b.run(nsteps, 1e-3, 0.9, 1e9, plain=True)
    dev.cuda.synchronize()
    I = (slice(None), slice(2, -2), slice(2, -2))
    for k in range(len(members)):
        assert np.array_equal(a.q(k)[I], b.q(k)[I]), k
    assert dev.equal(a.dt, b.dt)
    for k in (int(np.argmin(cells)), int(np.argmax(cells))):
        pt = members[k].patch
        print(f"[c2] member {k}: {pt.mx}x{pt.my} against the oracle", flush=True)
        qo, _, dto, _, _ = O.run(O.params_from(pt.params), pt.q, pt.aux, pt.dx, pt.dy, nsteps,
                                 1e-3, 0.9, 1e9)
        qg = a.q(k)
        assert np.array_equal(qg[I], qo[I]), k
        assert a.dt[k].item() == dto, k
        assert (qo[0][2:-2, 2:-2] > pt.params.dry_tolerance).sum() > 100
    print("[c2] all 256 members against the oracle (OpenMP over members)", flush=True)
    import os
    with O.libm("omp"):
        qos, dtos = O.run_patches([O.params_from(m.patch.params) for m in members],
                                  [m.patch.q for m in members], [m.patch.aux for m in members],
                                  [m.patch.dx for m in members], [m.patch.dy for m in members],
                                  nsteps, 1e-3, 0.9, 1e9, nthreads=min(16, os.cpu_count() or 1))
    dts = a.dt.cpu().numpy()
    for k in range(len(members)):
        assert np.array_equal(a.q(k)[I], qos[k][I]), k
        assert dts[k] == dtos[k], k


@pytest.mark.parametrize("bc,seg", [((1, 1, 1, 1), 8), ((3, 1, 3, 3), 2)])
def test_batched_resident_skips_dry_tiles(dev, bc, seg):
    """The resident batched split step (Ensemble's workspace) leaves each
    member's tiles whose neighbourhood stayed dry unread and unwritten, and
    its states and dt equal the same ensemble's without skipping (a workspace
    of the ordered step only) bit for bit over 160 steps, through a state
    written from outside in the middle (the bookkeeping starts over)."""
    from dclaw_amd import _lib
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    L = _lib.lib()
    members = [W.Member(W.fractal_sources(300, 200, dx=2.0, seed=5, nsrc=6), 0),
               W.Member(W.fractal_sources(130, 250, dx=2.0, seed=6, nsrc=3), 1)]
    members += _members(n=3, seed=11)
    print(f"[resident batched] bc {bc} seg {seg}", flush=True)
    a = S.Ensemble(members, bc=bc, seg_rows=seg)
    b = S.Ensemble(members, bc=bc, seg_rows=seg)
    nt = a.ntiles
    assert a.flags.numel() == L.dclaw_batched_resident_size(nt, len(members))
    b.flags = dev.zeros(2 * nt + 4, dtype=dev.int32, device="cuda")
    for e in (a, b):
        e.dt.fill_(1e-3)
    for n in range(80):
        for e in (a, b):
            e.step()
            e.cfl.zero_()
        if n == 40:
            for e in (a, b):
                e.current(1).copy_(e.current(1).clone())
            a.state_written()
    a.run(80, a.dt.clone(), 0.9, 1e9)
    b.run(80, b.dt.clone(), 0.9, 1e9)
    dev.cuda.synchronize()
    for k in range(len(members)):
        qa, qb = a.q(k), b.q(k)
        assert np.array_equal(qa, qb), f"member {k}: {(qa != qb).sum()} values differ"
    assert dev.equal(a.dt, b.dt)
    # member 0's far dry tiles are skipped: a marker put into the buffer the
    # next step writes survives that step
    npatch = len(members)
    age = a.flags[2 * nt + 8 + 5 * npatch:3 * nt + 8 + 5 * npatch].cpu().numpy()
    tiles = a.tiles.cpu().numpy().reshape(-1, 4)
    mine = np.flatnonzero(tiles[:, 0] == 0)
    old = mine[age[mine] >= 3]
    assert old.size > 0 and (age[mine] < 3).any()
    busy = tiles[mine[age[mine] < 3]]
    far = max(old, key=lambda t: min(abs(tiles[t, 1] - bx) * 58 + abs(tiles[t, 2] - bg) * seg
                                     for _, bx, bg, _ in busy))
    i = 1 + tiles[far, 1] * 58
    j = 1 + tiles[far, 2] * seg
    nxt = a.qa[0] if a.flip else a.qb[0]
    nxt[j + 1, i + 1, 6] = 12345.0
    a.step()
    dev.cuda.synchronize()
    assert a.current(0)[j + 1, i + 1, 6].item() == 12345.0


def test_fused_bc_equals_bound_then_step(dev):
    from dclaw_amd import step as S
    pt = slope_patch(90, 70, slope=0.8, phi_deg=25.0, seed=2)
    for bc in [(1, 1, 1, 1), (3, 3, 3, 3)]:
        r1 = S.PatchRunner(pt.q, pt.aux, pt.params, 1.0, 1.0, bc=bc)
        r2 = S.PatchRunner(pt.q, pt.aux, pt.params, 1.0, 1.0, bc=bc)
        r1.run(15, 1e-3, 0.9, 1e9, fused_bc=True)
        r2.run(15, 1e-3, 0.9, 1e9, fused_bc=False)
        dev.cuda.synchronize()
        assert np.array_equal(r1.q()[:, 2:-2, 2:-2], r2.q()[:, 2:-2, 2:-2])


def test_host_dropin_stepgrid(dev, oracle_mod):
    """stepgrid_dclaw_ (Fortran convention, host arrays) = the oracle's
    stepgrid on q and on aux(4..6)."""
    from dclaw_amd import _lib
    from dclaw_amd.params import to_c
    from helpers import advance_oracle
    O = oracle_mod
    L = _lib.lib()
    pt = slope_patch(75, 52, slope=0.7, phi_deg=26.0, seed=4)
    q, aux, dt = advance_oracle(O, pt, 6)
    cp = to_c(pt.params)
    _lib.check(L.dclaw_set_params(ctypes.byref(cp)))
    qh = np.array(q, order="F", copy=True)
    ah = np.array(aux, order="F", copy=True)
    dp = ctypes.POINTER(ctypes.c_double)
    ci = lambda v: ctypes.byref(ctypes.c_int(v))
    cd = lambda v: ctypes.byref(ctypes.c_double(v))
    cfl = ctypes.c_double(0.0)
    L.stepgrid_dclaw_(qh.ctypes.data_as(dp), ah.ctypes.data_as(dp), ci(2), ci(pt.mx), ci(pt.my),
                      ci(ah.shape[0]), cd(1.0), cd(1.0), cd(dt), ctypes.byref(cfl))
    qo, auxo, cflo = O.stepgrid(O.params_from(pt.params), q, aux, 1.0, 1.0, dt)
    assert rel_linf_wet(qh, qo, pt.params.dry_tolerance) <= TOL_STEP
    assert mask_equal(qh, qo, pt.params.dry_tolerance)
    assert np.allclose(ah, auxo, rtol=1e-12, atol=1e-300)
    assert abs(cfl.value - cflo) <= TOL_STEP * cflo


def test_host_dropin_stepgrid_concurrent_threads(dev, oracle_mod):
    """stepgrid_dclaw_ called from four threads at once (AMRClaw advanc's
    OpenMP loop over patches; ctypes releases the GIL during the call), each
    stepping its own patch three times: every result equals the same calls
    made one after another on one thread, bit for bit."""
    import threading

    from dclaw_amd import _lib
    from dclaw_amd.params import to_c
    from helpers import advance_oracle
    O = oracle_mod
    L = _lib.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    ci = lambda v: ctypes.byref(ctypes.c_int(v))
    cd = lambda v: ctypes.byref(ctypes.c_double(v))
    patches = []
    for k, (mx, my) in enumerate([(75, 52), (130, 41), (48, 96), (101, 77)]):
        pt = slope_patch(mx, my, slope=0.7, phi_deg=26.0, seed=10 + k)
        q, aux, dt = advance_oracle(O, pt, 4)
        patches.append((pt, q, aux, dt))
    cp = to_c(patches[0][0].params)
    _lib.check(L.dclaw_set_params(ctypes.byref(cp)))

    def run(k, out):
        pt, q, aux, dt = patches[k]
        qh = np.array(q, order="F", copy=True)
        ah = np.array(aux, order="F", copy=True)
        cfls = []
        for _ in range(3):
            c = ctypes.c_double(0.0)
            L.stepgrid_dclaw_(qh.ctypes.data_as(dp), ah.ctypes.data_as(dp), ci(2), ci(pt.mx),
                              ci(pt.my), ci(ah.shape[0]), cd(pt.dx), cd(pt.dy), cd(dt),
                              ctypes.byref(c))
            cfls.append(c.value)
        out[k] = (qh, ah, cfls)

    print("[host stepgrid threads] serial", flush=True)
    serial = {}
    for k in range(len(patches)):
        run(k, serial)
    print("[host stepgrid threads] 4 threads", flush=True)
    conc = {}
    ths = [threading.Thread(target=run, args=(k, conc)) for k in range(len(patches))]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    for k in range(len(patches)):
        assert np.array_equal(conc[k][0], serial[k][0]), k
        assert np.array_equal(conc[k][1], serial[k][1]), k
        assert conc[k][2] == serial[k][2], k


def test_host_dropin_stepgrid_with_fluxes(dev, oracle_mod):
    """stepgrid_fluxes_dclaw_: the same q, aux and cflgrid as stepgrid_dclaw_,
    and fm, fp, gm, gp equal to the oracle's step2 on the oracle's b4step2
    state bit for bit (what AMRClaw's fix-up reads after stepgrid)."""
    from dclaw_amd import _lib
    from dclaw_amd.params import to_c
    from helpers import advance_oracle
    O = oracle_mod
    L = _lib.lib()
    pt = slope_patch(61, 44, slope=0.7, phi_deg=26.0, seed=5)
    q, aux, dt = advance_oracle(O, pt, 6)
    cp = to_c(pt.params)
    _lib.check(L.dclaw_set_params(ctypes.byref(cp)))
    dp = ctypes.POINTER(ctypes.c_double)
    ci = lambda v: ctypes.byref(ctypes.c_int(v))
    cd = lambda v: ctypes.byref(ctypes.c_double(v))
    print("[host stepgrid with fluxes] 61x44", flush=True)
    q1, a1 = np.array(q, order="F", copy=True), np.array(aux, order="F", copy=True)
    c1 = ctypes.c_double(0.0)
    L.stepgrid_dclaw_(q1.ctypes.data_as(dp), a1.ctypes.data_as(dp), ci(2), ci(pt.mx), ci(pt.my),
                      ci(a1.shape[0]), cd(pt.dx), cd(pt.dy), cd(dt), ctypes.byref(c1))
    q2, a2 = np.array(q, order="F", copy=True), np.array(aux, order="F", copy=True)
    c2 = ctypes.c_double(0.0)
    f = [np.full_like(q2, np.nan, order="F") for _ in range(4)]
    L.stepgrid_fluxes_dclaw_(q2.ctypes.data_as(dp), a2.ctypes.data_as(dp), ci(2), ci(pt.mx),
                             ci(pt.my), ci(a2.shape[0]), cd(pt.dx), cd(pt.dy), cd(dt),
                             ctypes.byref(c2), *[a.ctypes.data_as(dp) for a in f])
    assert np.array_equal(q1, q2) and np.array_equal(a1, a2) and c1.value == c2.value
    P = O.params_from(pt.params)
    qb, ab = O.b4step2(P, q, aux, pt.dx, pt.dy)
    cflo, *fo = O.step2(P, qb, ab, pt.dx, pt.dy, dt)
    for k, (a, b) in enumerate(zip(f, fo)):
        assert np.array_equal(a, b), f"flux {k}"
    assert c2.value == cflo
