"""GPU parity: the fused gfx950 stepgrid against the CPU oracle.

Bars: per step relative L-inf over wet cells <= 1e-12, after 1000 steps
<= 1e-9, wet/dry mask identical; CFL numbers to 1e-12.  The GPU and the
oracle share the elementary functions (dcmath.h), the operation order and
the absence of FMA contraction, so in practice the results are bit-identical.
"""
import numpy as np
import pytest

from helpers import TOL_LONG, TOL_STEP, advance_oracle, mask_equal, rel_linf_wet, slope_patch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dclaw_amd import _lib
    _lib.lib()
    return torch


def _gpu_stepgrid(torch, patch, q, aux, dt):
    from dclaw_amd import step as S
    qi = S.to_device_q(q)
    qo = torch.zeros_like(qi)
    bed = S.bed_tensor(aux)
    prm = S.params_tensor([patch.params])
    dtd = torch.tensor([dt], dtype=torch.float64, device="cuda")
    cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
    S.stepgrid(qi, qo, bed, 1, patch.mx, patch.my, patch.dx, patch.dy, dtd, prm, cfl)
    torch.cuda.synchronize()
    return S.to_host_q(qo), float(cfl.item())


@pytest.mark.parametrize("nprior", [0, 5, 40])
def test_stepgrid_c0_one_step(dev, oracle_mod, nprior):
    from dclaw_amd import workloads as W
    O = oracle_mod
    pt = W.inclined_plane_mound()
    q, aux, dt = advance_oracle(O, pt, nprior)
    print(f"[c0 one step] prior steps {nprior} dt {dt}", flush=True)
    qg, cflg = _gpu_stepgrid(dev, pt, q, aux, dt)
    qo, auxo, cflo = O.stepgrid(O.params_from(pt.params), q, aux, pt.dx, pt.dy, dt)
    err = rel_linf_wet(qg[:, 2:-2, 2:-2], qo[:, 2:-2, 2:-2], pt.params.dry_tolerance)
    print("rel err", err, "cfl", cflg, cflo)
    assert mask_equal(qg, qo, pt.params.dry_tolerance)
    assert err <= TOL_STEP
    assert abs(cflg - cflo) <= TOL_STEP * cflo
    # ghost cells carry the b4step2-adjusted state, as the reference leaves them
    assert rel_linf_wet(qg, qo, pt.params.dry_tolerance) <= TOL_STEP


@pytest.mark.parametrize("mx,my,bc", [(130, 70, (3, 3, 3, 3)), (59, 33, (1, 3, 3, 1)),
                                      (1, 1, (1, 1, 1, 1)), (7, 90, (3, 1, 1, 3))])
def test_stepgrid_ragged_walls(dev, oracle_mod, mx, my, bc):
    O = oracle_mod
    pt = slope_patch(mx, my, slope=0.8, phi_deg=25.0, seed=mx + my)
    q, aux, dt = advance_oracle(O, pt, 8, bc=bc)
    print(f"[ragged] {mx}x{my} bc {bc}", flush=True)
    qg, cflg = _gpu_stepgrid(dev, pt, q, aux, dt)
    qo, _, cflo = O.stepgrid(O.params_from(pt.params), q, aux, pt.dx, pt.dy, dt)
    assert mask_equal(qg, qo, pt.params.dry_tolerance)
    assert rel_linf_wet(qg, qo, pt.params.dry_tolerance) <= TOL_STEP
    assert abs(cflg - cflo) <= TOL_STEP * max(cflo, 1e-300)


def test_dry_patch(dev, oracle_mod):
    O = oracle_mod
    pt = slope_patch(80, 40, h0=0.0)
    q = O.bc_fill(pt.q, (1, 1, 1, 1))
    qg, cflg = _gpu_stepgrid(dev, pt, q, pt.aux, 0.01)
    qo, _, cflo = O.stepgrid(O.params_from(pt.params), q, pt.aux, pt.dx, pt.dy, 0.01)
    assert np.array_equal(qg, qo)
    assert cflg == 0.0 and cflo == 0.0


def test_c0_long_run(dev, oracle_mod):
    """1000 steps of the 200x200 inclined-plane workload, device time loop
    (bound, stepgrid, device dt rule) against the oracle's loop."""
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    O = oracle_mod
    pt = W.inclined_plane_mound()
    nsteps = 1000
    print("[c0 long] oracle run", flush=True)
    qo, _, dto, cflo, dtso = O.run(O.params_from(pt.params), pt.q, pt.aux, pt.dx, pt.dy, nsteps,
                                   1e-3, 0.9, 1e9)
    print("[c0 long] gpu run", flush=True)
    r = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy)
    hist = dev.zeros(nsteps, dtype=dev.float64, device="cuda")
    r.run(nsteps, 1e-3, 0.9, 1e9, cfl_hist=hist)
    dev.cuda.synchronize()
    qg = r.q()
    cflg = hist.cpu().numpy()
    err = rel_linf_wet(qg[:, 2:-2, 2:-2], qo[:, 2:-2, 2:-2], pt.params.dry_tolerance)
    print("rel err after", nsteps, "steps:", err, "cfl rel", np.max(np.abs(cflg - cflo) / cflo))
    assert mask_equal(qg[:, 2:-2, 2:-2], qo[:, 2:-2, 2:-2], pt.params.dry_tolerance)
    assert err <= TOL_LONG
    assert np.max(np.abs(cflg - cflo) / cflo) <= TOL_LONG
