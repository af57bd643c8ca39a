"""GPU: the AMR level transfers (dclaw_amr_average_dev, dclaw_amr_interp_dev)
against the oracle's restatements of update.f90 and filval.f90, bit for bit,
on random wet/dry patches; and the coarse -> fine -> coarse round trip."""
import numpy as np
import pytest

from test_amr_pins import DRY, GRAV, RHO_F, _random_levels

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dclaw_amd import _lib
    _lib.lib()
    return torch


def _to_dev(torch, a, dtype=None):
    """Fortran (c, nx, ny) array -> device tensor (ny, nx, c) of the same bytes."""
    a = np.asfortranarray(a)
    t = torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0) if a.ndim == 3 else a.T))
    return t.to("cuda", dtype=dtype or t.dtype)


def _to_host(t):
    a = t.cpu().numpy()
    return np.asfortranarray(a.transpose(2, 1, 0) if a.ndim == 3 else a.T)


@pytest.mark.parametrize("seed,mic,mjc,rx,ry", [(0, 9, 7, 2, 2), (1, 60, 45, 4, 4),
                                                (2, 33, 70, 4, 2), (3, 17, 12, 3, 5),
                                                (4, 130, 90, 2, 2)])
def test_amr_transfers_match_oracle(dev, oracle_mod, seed, mic, mjc, rx, ry):
    from dclaw_amd import amr
    O = oracle_mod
    torch = dev
    print(f"[amr] seed {seed} coarse {mic}x{mjc} ratio {rx}x{ry}", flush=True)
    rng = np.random.default_rng(seed)
    valc, auxc, val, auxf, ns = _random_levels(rng, mic, mjc, rx, ry)
    sea = 0.5
    ref = O.amr_interp(valc, auxc, val, auxf, ns, rx, ry, DRY, sea, RHO_F, GRAV)
    vc, ac, vf, af = (_to_dev(torch, x) for x in (valc, auxc, val, auxf))
    nsd = _to_dev(torch, np.asfortranarray(ns, dtype=np.uint8))
    amr.interpolate_to_fine(vc, ac, vf, af, nsd, rx, ry, DRY, sea, RHO_F, GRAV)
    torch.cuda.synchronize()
    got = _to_host(vf)
    assert np.array_equal(got, ref, equal_nan=True), f"{(got != ref).sum()} values differ"
    # fine -> coarse over most of the coarse patch
    ni, nj, ci0, cj0, fi0, fj0 = mic - 2, mjc - 3, 1, 2, 2, 2
    refc = O.amr_average(valc, auxc, ref, auxf, ci0, cj0, ni, nj, fi0, fj0, rx, ry, DRY)
    amr.average_to_coarse(vc, ac, vf, af, ci0, cj0, ni, nj, fi0, fj0, rx, ry, DRY)
    torch.cuda.synchronize()
    gotc = _to_host(vc)
    assert np.array_equal(gotc, refc), f"{(gotc != refc).sum()} values differ"


def test_amr_round_trip_wet_flat(dev):
    """All wet on a flat bed: interpolating to the fine level and averaging
    back returns the coarse depth and momenta (mass and momentum of each
    block are conserved by both transfers there)."""
    from dclaw_amd import amr
    torch = dev
    rng = np.random.default_rng(9)
    mic, mjc, r = 40, 30, 4
    valc = np.zeros((7, mic, mjc), order="F")
    valc[0] = rng.uniform(1.0, 3.0, (mic, mjc))
    valc[1] = valc[0] * rng.uniform(-0.5, 0.5, (mic, mjc))
    valc[2] = valc[0] * rng.uniform(-0.5, 0.5, (mic, mjc))
    auxc = np.zeros((1, mic, mjc), order="F")
    mit, mjt = (mic - 2) * r + 4, (mjc - 2) * r + 4
    vc, ac = _to_dev(torch, valc), _to_dev(torch, auxc)
    vf = torch.zeros((mjt, mit, 7), dtype=torch.float64, device="cuda")
    af = torch.zeros((mjt, mit, 1), dtype=torch.float64, device="cuda")
    ns = torch.ones((mjt, mit), dtype=torch.uint8, device="cuda")
    amr.interpolate_to_fine(vc, ac, vf, af, ns, r, r, DRY, 0.0, RHO_F, GRAV)
    back = vc.clone()
    amr.average_to_coarse(back, ac, vf, af, 1, 1, mic - 2, mjc - 2, 2, 2, r, r, DRY)
    torch.cuda.synchronize()
    a, b = _to_host(back)[:3, 1:-1, 1:-1], valc[:3, 1:-1, 1:-1]
    np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-13)


def test_amr_argument_checks(dev):
    from dclaw_amd import _lib, amr
    torch = dev
    q = torch.zeros((10, 10, 7), dtype=torch.float64, device="cuda")
    a = torch.zeros((10, 10, 1), dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.DClawError, match="outside"):
        amr.average_to_coarse(q, a, q, a, 0, 0, 4, 4, 0, 0, 4, 4, DRY)   # 16 fine cells > 10
    ns = torch.ones((10, 10), dtype=torch.uint8, device="cuda")
    with pytest.raises(_lib.DClawError, match="smaller"):
        amr.interpolate_to_fine(q, a, q, a, ns, 2, 2, DRY, 0.0, RHO_F, GRAV)
