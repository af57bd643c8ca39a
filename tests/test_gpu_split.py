"""GPU: the split step (strip kernel without src2, then src2 over the tiles
it flagged) against the fused step, single patch and batched, bit for bit."""
import numpy as np
import pytest

from helpers import advance_oracle, slope_patch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dclaw_amd import _lib
    _lib.lib()
    return torch


@pytest.mark.parametrize("mx,my,seg,bc", [(130, 70, 8, (3, 3, 3, 3)), (59, 33, 4, (1, 3, 3, 1)),
                                          (200, 150, 16, (1, 1, 1, 1)), (7, 90, 32, None)])
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
    assert set(fl.tolist()) <= {0, 1} and fl.sum() > 0


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
    b = S.Ensemble(members, seg_rows=16, split_src2=True)
    a.run(20, 1e-3, 0.9, 1e9)
    b.run(20, 1e-3, 0.9, 1e9)
    dev.cuda.synchronize()
    for k in range(len(members)):
        assert np.array_equal(a.q(k), b.q(k)), k
    assert dev.equal(a.dt, b.dt)


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
