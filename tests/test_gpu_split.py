"""GPU: the split step (strip kernel without src2, then src2 over the tiles
it flagged) against the fused step, single patch and batched, bit for bit."""
import numpy as np
import pytest

from helpers import TOL_STEP, advance_oracle, mask_equal, rel_linf_wet, slope_patch

pytestmark = pytest.mark.gpu


def valid_flags(f) -> bool:
    """Tile flags: 0, or bit 0 (a cell needs src2) with any of the four band
    bits."""
    f = np.asarray(f)
    return bool(np.all((f == 0) | ((f & 1) == 1)) and np.all((f >= 0) & (f < 32)))


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dclaw_amd import _lib
    _lib.lib()
    return torch


@pytest.mark.parametrize("mx,my,seg,bc", [(130, 70, 8, (3, 3, 3, 3)), (59, 33, 4, (1, 3, 3, 1)),
                                          (200, 150, 16, (1, 1, 1, 1)), (7, 90, 32, None),
                                          (200, 40, 1, (1, 1, 1, 1)), (90, 37, 2, (3, 1, 3, 1))])
def test_split_equals_fused(dev, oracle_mod, mx, my, seg, bc):
    from dclaw_amd import _lib
    from dclaw_amd import step as S
    O = oracle_mod
    pt = slope_patch(mx, my, slope=0.8, phi_deg=25.0, seed=mx + my)
    q, aux, dt = advance_oracle(O, pt, 6, bc=bc or (1, 1, 1, 1))
    print(f"[split] {mx}x{my} seg {seg} bc {bc}", flush=True)
    L = _lib.lib()
    prev = L.dclaw_set_seg_rows(seg)
    try:
        pdev = S.params_tensor([pt.params])
        dtd = dev.tensor([dt], dtype=dev.float64, device="cuda")
        bed = S.bed_tensor(aux)
        qi = S.to_device_q(q)
        qf, qs = dev.zeros_like(qi), dev.zeros_like(qi)
        cf = dev.zeros(1, dtype=dev.float64, device="cuda")
        cs = dev.zeros(1, dtype=dev.float64, device="cuda")
        if bc is None:
            S.stepgrid(qi, qf, bed, 1, mx, my, pt.dx, pt.dy, dtd, pdev, cf)
        else:
            S.stepgrid_bc(qi, qf, bed, 1, mx, my, pt.dx, pt.dy, bc, dtd, pdev, cf)
        flags = dev.full((L.dclaw_tile_count(mx, my),), -1, dtype=dev.int32, device="cuda")
        S.stepgrid_split(qi, qs, bed, 1, mx, my, pt.dx, pt.dy, bc, dtd, pdev, cs, flags)
        dev.cuda.synchronize()
    finally:
        L.dclaw_set_seg_rows(prev)
    a, b = S.to_host_q(qs), S.to_host_q(qf)
    assert np.array_equal(a, b), f"{(a != b).sum()} values differ"
    assert cs.item() == cf.item()
    fl = flags.cpu().numpy()
    assert valid_flags(fl) and (fl > 0).any()


def test_split_runner_matches_oracle_long(dev, oracle_mod):
    """200 steps of the split runner (device dt rule) against the oracle."""
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    from helpers import TOL_LONG, mask_equal, rel_linf_wet
    O = oracle_mod
    pt = W.inclined_plane_mound()
    nsteps = 200
    r = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy, split_src2=True)
    r.run(nsteps, 1e-3, 0.9, 1e9)
    qo, _, dto, _, _ = O.run(O.params_from(pt.params), pt.q, pt.aux, pt.dx, pt.dy, nsteps, 1e-3,
                             0.9, 1e9)
    qg = r.q()
    I = (slice(None), slice(2, -2), slice(2, -2))
    assert mask_equal(qg[I], qo[I], pt.params.dry_tolerance)
    assert rel_linf_wet(qg[I], qo[I], pt.params.dry_tolerance) <= TOL_LONG
    assert r.dt.item() == pytest.approx(dto, rel=1e-12)


def test_split_batched_equals_fused(dev):
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    rng = np.random.default_rng(3)
    members = []
    for k in range(5):
        mx, my = int(rng.integers(10, 150)), int(rng.integers(10, 150))
        members.append(W.Member(slope_patch(mx, my, slope=0.7, phi_deg=25.0, seed=k), seed=k))
    a = S.Ensemble(members, seg_rows=16, split_src2=False)
    b = S.Ensemble(members, seg_rows=16, split_src2=True)                 # wet-first order
    c = S.Ensemble(members, seg_rows=16, split_src2=True, ordered=False)  # work-list order
    a.run(20, 1e-3, 0.9, 1e9)
    b.run(20, 1e-3, 0.9, 1e9)
    c.run(20, 1e-3, 0.9, 1e9)
    dev.cuda.synchronize()
    for k in range(len(members)):
        assert np.array_equal(a.q(k), b.q(k)), k
        assert np.array_equal(a.q(k), c.q(k)), k
    assert dev.equal(a.dt, b.dt) and dev.equal(a.dt, c.dt)
    nt = b.ntiles
    order = b.flags[nt:2 * nt].cpu().numpy()
    assert np.array_equal(np.sort(order), np.arange(nt))


def _ent_patch():
    from dclaw_amd import workloads as W
    pt = slope_patch(90, 50, slope=0.9, phi_deg=22.0, seed=12)
    rng = np.random.default_rng(3)
    pe = W.make_patch(np.zeros((90, 50)), np.zeros((90, 50)), 0.6, 0.8, pt.params, 1.0, 1.0,
                      maux=7, h_e=rng.uniform(0.0, 0.6, (90, 50)))
    pt.aux = np.concatenate([pt.aux, pe.aux[6:7]], axis=0)
    pt.params.entrainment, pt.params.entrainment_method = 1, 0
    pt.params.entrainment_rate = 0.3
    return pt


def test_entrainment_runner_matches_oracle(dev, oracle_mod):
    """Entrainment (erodible layer in aux(7)) through the device loop, split
    and fused, against the oracle over 60 steps."""
    from dclaw_amd import step as S
    from helpers import TOL_LONG, mask_equal, rel_linf_wet
    O = oracle_mod
    pt = _ent_patch()
    n = 60
    qo, _, dto, _, _ = O.run(O.params_from(pt.params), pt.q, pt.aux, pt.dx, pt.dy, n, 1e-3, 0.9,
                             1e9)
    I = (slice(None), slice(2, -2), slice(2, -2))
    for split in (True, False):
        r = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy, split_src2=split)
        assert r.maux == 7
        r.run(n, 1e-3, 0.9, 1e9)
        qg = r.q()
        assert mask_equal(qg[I], qo[I], pt.params.dry_tolerance)
        assert rel_linf_wet(qg[I], qo[I], pt.params.dry_tolerance) <= TOL_LONG
        assert np.array_equal(qg[6][2:-2, 2:-2], qo[6][2:-2, 2:-2])   # eroded depth
    assert (qo[6] > 0).any()


def test_entrainment_in_a_batch(dev, oracle_mod):
    """A batch mixing an entrainment member with plain members."""
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    O = oracle_mod
    members = [W.Member(_ent_patch(), 0), W.Member(slope_patch(40, 60, seed=4), 1)]
    ens = S.Ensemble(members, seg_rows=16)
    ens.run(20, 1e-3, 0.9, 1e9)
    for k, m in enumerate(members):
        pt = m.patch
        qo, _, _, _, _ = O.run(O.params_from(pt.params), pt.q, pt.aux, pt.dx, pt.dy, 20, 1e-3,
                               0.9, 1e9)
        I = (slice(None), slice(2, -2), slice(2, -2))
        assert np.array_equal(ens.q(k)[I], qo[I]), k


@pytest.mark.parametrize("mx,my,seg,dry_rows,bc", [
    (130, 70, 8, 32, (3, 3, 3, 3)), (59, 33, 4, 20, (1, 3, 3, 1)),
    (200, 150, 16, 48, (1, 1, 1, 1)), (7, 90, 32, 48, None),
    (600, 500, 6, 48, (1, 1, 1, 1)), (300, 300, 1, 200, (1, 3, 1, 3)),
])
def test_dry_pass_equals_fused(dev, oracle_mod, mx, my, seg, dry_rows, bc):
    """The split step with its dry pass (classify kernel, strip kernel over the
    wet list, src2 over the list) equals the fused step on every cell, ghosts
    included; q_out starts as NaN so an unwritten cell shows."""
    from dclaw_amd import _lib
    from dclaw_amd import step as S
    O = oracle_mod
    pt = slope_patch(mx, my, slope=0.8, phi_deg=25.0, seed=mx + my)
    q, aux, dt = advance_oracle(O, pt, 6, bc=bc or (1, 1, 1, 1))
    print(f"[dry pass] {mx}x{my} seg {seg} dry rows {dry_rows} bc {bc}", flush=True)
    L = _lib.lib()
    prev = L.dclaw_set_seg_rows(seg)
    prev_dry = L.dclaw_set_dry_rows(dry_rows)
    try:
        ntiles = L.dclaw_tile_count(mx, my)
        nws = L.dclaw_split_workspace_size(mx, my)
        assert nws == 3 * ntiles + 4
        pdev = S.params_tensor([pt.params])
        dtd = dev.tensor([dt], dtype=dev.float64, device="cuda")
        bed = S.bed_tensor(aux)
        qi = S.to_device_q(q)
        qf = dev.zeros_like(qi)
        qs = dev.full_like(qi, float("nan"))
        cf = dev.zeros(1, dtype=dev.float64, device="cuda")
        cs = dev.zeros(1, dtype=dev.float64, device="cuda")
        if bc is None:
            S.stepgrid(qi, qf, bed, 1, mx, my, pt.dx, pt.dy, dtd, pdev, cf)
        else:
            S.stepgrid_bc(qi, qf, bed, 1, mx, my, pt.dx, pt.dy, bc, dtd, pdev, cf)
        ws = dev.full((nws,), -1, dtype=dev.int32, device="cuda")
        S.stepgrid_split(qi, qs, bed, 1, mx, my, pt.dx, pt.dy, bc, dtd, pdev, cs, ws, mode="dry")
        dev.cuda.synchronize()
    finally:
        L.dclaw_set_seg_rows(prev)
        L.dclaw_set_dry_rows(prev_dry)
    a, b = S.to_host_q(qs), S.to_host_q(qf)
    assert not np.isnan(a).any(), f"{np.isnan(a).sum()} values not written"
    assert np.array_equal(a, b), f"{(a != b).sum()} values differ"
    assert cs.item() == cf.item()
    w = ws.cpu().numpy()
    nwet = int(w[2 * ntiles])
    assert 0 < nwet <= ntiles
    assert w[2 * ntiles + 1] >= nwet                      # the work cursor passed the list
    assert valid_flags(w[:ntiles])
    listed = w[ntiles:ntiles + nwet]
    assert len(set(listed.tolist())) == nwet and listed.min() >= 0 and listed.max() < ntiles
    # unlisted tiles are dry: flag 0
    unl = np.setdiff1d(np.arange(ntiles), listed)
    assert (w[unl] == 0).all()
    if mx * my >= 600 * 500:
        assert nwet < ntiles          # the mound's far tail leaves some tiles dry


def test_dry_pass_runner_long(dev):
    """150 steps of the runner on a 700 x 520 fractal patch with 64 sources
    in each split mode (dry pass, one launch in tile order, one launch with
    the previous step's wet tiles first, the same with the dry segments in
    tall groups): identical states and dt."""
    from dclaw_amd import _lib
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    pt = W.fractal_sources(700, 520, dx=2.0, seed=3)
    L = _lib.lib()
    prev = L.dclaw_set_seg_rows(8)
    try:
        ra = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy, split_mode="dry")
        rb = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy, split_mode="single")
        rc = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy, split_mode="ordered")
        rd = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy, split_mode="hybrid")
        ra.run(150, 1e-3, 0.9, 1e9)
        rb.run(150, 1e-3, 0.9, 1e9)
        rc.run(150, 1e-3, 0.9, 1e9)
        rd.run(150, 1e-3, 0.9, 1e9)
        dev.cuda.synchronize()
        ntiles = L.dclaw_tile_count(700, 520)
        nwet = int(ra.flags[2 * ntiles].item())
        order = rc.flags[ntiles:2 * ntiles].cpu().numpy()
        nfirst = int(rc.flags[2 * ntiles + 2].item())
    finally:
        L.dclaw_set_seg_rows(prev)
    qa, qb, qc, qd = ra.q(), rb.q(), rc.q(), rd.q()
    assert np.array_equal(qa, qb), f"{(qa != qb).sum()} values differ"
    assert np.array_equal(qa, qc), f"{(qa != qc).sum()} values differ"
    assert np.array_equal(qa, qd), f"{(qa != qd).sum()} values differ"
    assert ra.dt.item() == rb.dt.item() == rc.dt.item() == rd.dt.item()
    assert 0 < nwet < ntiles
    # the last step's order: a permutation of the tiles, flagged ones first
    assert np.array_equal(np.sort(order), np.arange(ntiles))
    assert 0 < nfirst < ntiles


def test_tuner_restores_state(dev):
    """The autotuner times (height, dry pass) candidates on copies of the
    runner's state and leaves state, dt and the library's height as chosen."""
    from dclaw_amd import _lib, autotune
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    pt = W.fractal_sources(300, 200, dx=2.0, seed=4, nsrc=8)
    L = _lib.lib()
    prev = L.dclaw_set_seg_rows(32)
    try:
        r = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy)
        r.set_dt(1e-3)
        r.step()
        q0, dt0 = r.q(), r.dt.item()
        h, times = autotune.tune_patch(r, candidates=(4, 8), reps=2)
        assert set(times) == {(h_, m) for h_ in (4, 8)
                              for m in ("single", "ordered", "dry", "hybrid")}
        assert (h, r.split_mode) == min(times, key=times.get)
        assert L.dclaw_set_seg_rows(h) == h
        assert np.array_equal(r.q(), q0) and r.dt.item() == dt0
    finally:
        L.dclaw_set_seg_rows(prev)


@pytest.mark.parametrize("mx,my,seg,bc", [(130, 70, 8, (3, 3, 3, 3)), (59, 33, 4, (1, 3, 3, 1)),
                                          (600, 500, 6, (1, 1, 1, 1)), (7, 90, 32, None),
                                          (200, 40, 1, (1, 1, 1, 1)), (90, 37, 2, (3, 1, 3, 1))])
def test_ordered_equals_fused(dev, oracle_mod, mx, my, seg, bc):
    """The split step with the wet-first tile order, starting from arbitrary
    flags in the workspace, equals the fused step on every cell; the order it
    used is a permutation of the tiles with the previously flagged first."""
    from dclaw_amd import _lib
    from dclaw_amd import step as S
    O = oracle_mod
    pt = slope_patch(mx, my, slope=0.8, phi_deg=25.0, seed=mx + my)
    q, aux, dt = advance_oracle(O, pt, 6, bc=bc or (1, 1, 1, 1))
    print(f"[ordered] {mx}x{my} seg {seg} bc {bc}", flush=True)
    L = _lib.lib()
    prev = L.dclaw_set_seg_rows(seg)
    try:
        ntiles = L.dclaw_tile_count(mx, my)
        nws = L.dclaw_split_workspace_size(mx, my)
        rng = np.random.default_rng(mx)
        prev_flags = (rng.random(ntiles) < 0.3).astype(np.int32)
        ws = dev.full((nws,), -7, dtype=dev.int32, device="cuda")
        ws[:ntiles] = dev.from_numpy(prev_flags).to("cuda")
        pdev = S.params_tensor([pt.params])
        dtd = dev.tensor([dt], dtype=dev.float64, device="cuda")
        bed = S.bed_tensor(aux)
        qi = S.to_device_q(q)
        qf = dev.zeros_like(qi)
        qs = dev.full_like(qi, float("nan"))
        cf = dev.zeros(1, dtype=dev.float64, device="cuda")
        cs = dev.zeros(1, dtype=dev.float64, device="cuda")
        if bc is None:
            S.stepgrid(qi, qf, bed, 1, mx, my, pt.dx, pt.dy, dtd, pdev, cf)
        else:
            S.stepgrid_bc(qi, qf, bed, 1, mx, my, pt.dx, pt.dy, bc, dtd, pdev, cf)
        S.stepgrid_split(qi, qs, bed, 1, mx, my, pt.dx, pt.dy, bc, dtd, pdev, cs, ws,
                         mode="ordered")
        dev.cuda.synchronize()
    finally:
        L.dclaw_set_seg_rows(prev)
    a, b = S.to_host_q(qs), S.to_host_q(qf)
    assert not np.isnan(a).any()
    assert np.array_equal(a, b), f"{(a != b).sum()} values differ"
    assert cs.item() == cf.item()
    w = ws.cpu().numpy()
    order = w[ntiles:2 * ntiles]
    nfirst = int(w[2 * ntiles + 2])
    assert np.array_equal(np.sort(order), np.arange(ntiles))
    assert nfirst == int(prev_flags.sum())
    assert set(order[:nfirst].tolist()) == set(np.flatnonzero(prev_flags).tolist())
    assert valid_flags(w[:ntiles])


@pytest.mark.parametrize("mx,my,seg,bc", [(130, 70, 8, (3, 3, 3, 3)), (59, 33, 4, (1, 3, 3, 1)),
                                          (600, 500, 6, (1, 1, 1, 1)), (7, 90, 5, None),
                                          (200, 40, 1, (1, 1, 1, 1)), (90, 37, 2, (3, 1, 3, 1))])
def test_hybrid_equals_fused(dev, oracle_mod, mx, my, seg, bc):
    """The hybrid split step (the previous step's flagged 4-segment groups
    as single segments first, the others as one tall segment each), from
    arbitrary flags, equals the fused step on every cell; its work list holds
    every segment of the flagged groups and every other group once; a group's
    new src2 flag is on each of its segments."""
    from dclaw_amd import _lib
    from dclaw_amd import step as S
    O = oracle_mod
    pt = slope_patch(mx, my, slope=0.8, phi_deg=25.0, seed=mx + my + 1)
    q, aux, dt = advance_oracle(O, pt, 6, bc=bc or (1, 1, 1, 1))
    print(f"[hybrid] {mx}x{my} seg {seg} bc {bc}", flush=True)
    L = _lib.lib()
    prev = L.dclaw_set_seg_rows(seg)
    try:
        ntiles = L.dclaw_tile_count(mx, my)
        nws = L.dclaw_split_workspace_size(mx, my)
        nstrips = -(-mx // 58)
        nseg = ntiles // nstrips
        rng = np.random.default_rng(mx + 1)
        prev_flags = (rng.random(ntiles) < 0.2).astype(np.int32)
        ws = dev.full((nws,), -7, dtype=dev.int32, device="cuda")
        ws[:ntiles] = dev.from_numpy(prev_flags).to("cuda")
        pdev = S.params_tensor([pt.params])
        dtd = dev.tensor([dt], dtype=dev.float64, device="cuda")
        bed = S.bed_tensor(aux)
        qi = S.to_device_q(q)
        qf = dev.zeros_like(qi)
        qs = dev.full_like(qi, float("nan"))
        cf = dev.zeros(1, dtype=dev.float64, device="cuda")
        cs = dev.zeros(1, dtype=dev.float64, device="cuda")
        if bc is None:
            S.stepgrid(qi, qf, bed, 1, mx, my, pt.dx, pt.dy, dtd, pdev, cf)
        else:
            S.stepgrid_bc(qi, qf, bed, 1, mx, my, pt.dx, pt.dy, bc, dtd, pdev, cf)
        S.stepgrid_split(qi, qs, bed, 1, mx, my, pt.dx, pt.dy, bc, dtd, pdev, cs, ws,
                         mode="hybrid")
        dev.cuda.synchronize()
    finally:
        L.dclaw_set_seg_rows(prev)
    a, b = S.to_host_q(qs), S.to_host_q(qf)
    assert not np.isnan(a).any()
    assert np.array_equal(a, b), f"{(a != b).sum()} values differ"
    assert cs.item() == cf.item()
    w = ws.cpu().numpy()
    nw, nd = int(w[3 * ntiles]), int(w[3 * ntiles + 1])
    wet_list = w[ntiles:ntiles + nw]
    dry_list = w[2 * ntiles:2 * ntiles + nd]
    ngi = -(-nseg // 4)
    flagged = set()
    for g in range(nstrips * ngi):
        st, gi = g % nstrips, g // nstrips
        segs = [s_ * nstrips + st for s_ in range(gi * 4, min(nseg, gi * 4 + 4))]
        if prev_flags[segs].any():
            flagged.update(segs)
    assert set(wet_list.tolist()) == flagged and len(wet_list) == len(flagged)
    assert len(dry_list) + len({t // nstrips // 4 * nstrips + t % nstrips for t in flagged}) \
        == nstrips * ngi
    assert valid_flags(w[:ntiles])


@pytest.mark.parametrize("mode", ["single", "ordered", "dry", "hybrid"])
@pytest.mark.parametrize("seg", [1, 2])
def test_short_segments_match_oracle(dev, oracle_mod, mode, seg):
    """One- and two-row segments (the autotuner's pick on narrow patches,
    where few strips leave the launch short of work) give the oracle's step
    in every split mode, on a 200 x 40 slope patch; the second call runs with
    the flags the first one left."""
    from dclaw_amd import _lib
    from dclaw_amd import step as S
    O = oracle_mod
    mx, my = 200, 40
    pt = slope_patch(mx, my, slope=0.8, phi_deg=25.0, seed=7)
    q, aux, dt = advance_oracle(O, pt, 8, bc=(1, 1, 1, 1))
    print(f"[short segments] {mode} seg {seg}", flush=True)
    L = _lib.lib()
    prev = L.dclaw_set_seg_rows(seg)
    try:
        ws = dev.zeros((L.dclaw_split_workspace_size(mx, my),), dtype=dev.int32, device="cuda")
        pdev = S.params_tensor([pt.params])
        dtd = dev.tensor([dt], dtype=dev.float64, device="cuda")
        bed = S.bed_tensor(aux)
        qi = S.to_device_q(q)
        qs = dev.full_like(qi, float("nan"))
        cs = dev.zeros(1, dtype=dev.float64, device="cuda")
        for _ in range(2):
            S.stepgrid_split(qi, qs, bed, 1, mx, my, pt.dx, pt.dy, None, dtd, pdev, cs, ws,
                             mode=mode)
        dev.cuda.synchronize()
    finally:
        L.dclaw_set_seg_rows(prev)
    qg = S.to_host_q(qs)
    qo, _, cflo = O.stepgrid(O.params_from(pt.params), q, aux, pt.dx, pt.dy, dt)
    assert not np.isnan(qg).any()
    assert mask_equal(qg, qo, pt.params.dry_tolerance)
    assert rel_linf_wet(qg, qo, pt.params.dry_tolerance) <= TOL_STEP
    assert abs(cs.item() - cflo) <= TOL_STEP * cflo


@pytest.mark.parametrize("seg,bc", [(8, (1, 1, 1, 1)), (2, (3, 1, 3, 3)), (1, (1, 3, 1, 3))])
def test_resident_split_skips_dry_tiles(dev, seg, bc):
    """The resident wet-first split step (PatchRunner's workspace) leaves the
    tiles whose neighbourhood stayed dry unread and unwritten, and its states
    and dt equal the same loop's without skipping (a workspace too small for
    the dry-tile ages) bit for bit over 160 steps, through a tuner run in the
    middle (which rewrites the state and must restart the bookkeeping)."""
    from dclaw_amd import _lib, autotune
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    pt = W.fractal_sources(300, 200, dx=2.0, seed=5, nsrc=6)
    L = _lib.lib()
    prev = L.dclaw_set_seg_rows(seg)
    try:
        print(f"[resident] seg {seg} bc {bc}", flush=True)
        ra = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy, bc=bc, split_mode="ordered")
        rb = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy, bc=bc, split_mode="ordered")
        ntiles = L.dclaw_tile_count(300, 200)
        rb.flags = dev.zeros(L.dclaw_split_workspace_size(300, 200), dtype=dev.int32,
                             device="cuda")
        assert ra.flags.numel() >= L.dclaw_split_resident_size(300, 200)
        for r in (ra, rb):
            r.set_dt(1e-3)
        for n in range(80):
            ra.step()
            rb.step()
            if n == 40:
                h, _ = autotune.tune_patch(ra, candidates=(seg, 2 * seg), reps=2)
                L.dclaw_set_seg_rows(seg)
                ra.split_mode = "ordered"
        ra.run(80, ra.dt.clone(), 0.9, 1e9)
        rb.run(80, rb.dt.clone(), 0.9, 1e9)
        dev.cuda.synchronize()
        assert np.array_equal(ra.q(), rb.q()), f"{(ra.q() != rb.q()).sum()} values differ"
        assert ra.dt.item() == rb.dt.item()
        # a dry corner tile far from every source is skipped: a marker put into
        # the buffer the next step writes survives that step
        age = ra.flags[3 * ntiles + 12:4 * ntiles + 12].cpu().numpy()
        assert (age >= 3).any() and (age < 3).any()
        nstrips = -(-300 // L.dclaw_strip_width())
        # the skipped tile farthest from any tile that is not
        a2 = age.reshape(-1, nstrips)
        busy = np.argwhere(a2 < 3)
        far = [(min(max(abs(s - bs) * seg, abs(x - bx) * 58) for bs, bx in busy), s * nstrips + x)
               for s, x in np.argwhere(a2 >= 3)]
        t = max(far)[1]
        i = 1 + (t % nstrips) * L.dclaw_strip_width()
        j = 1 + (t // nstrips) * seg
        ra.qb[j + 1, i + 1, 6] = 12345.0
        ra.step()
        dev.cuda.synchronize()
        assert ra.qa[j + 1, i + 1, 6].item() == 12345.0
    finally:
        L.dclaw_set_seg_rows(prev)
