"""GPU checks at the full size of the 4096 x 4096 workload (the bench's):

* the first two steps of the device time loop against the oracle's (every
  interior cell: wet/dry mask identical, state within 1e-12 relative, CFL);
* the run does not depend on the tiling (segment height and split mode) bit
  for bit;
* a lake at rest (flat surface over the whole fractal bed, every tile wet,
  hydrostatic pore pressure) stays at rest: the step is well balanced over
  the whole patch, across every strip and segment boundary.
"""
import numpy as np
import pytest

from helpers import TOL_STEP, mask_equal, rel_linf_wet

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dclaw_amd import _lib
    _lib.lib()
    return torch


@pytest.fixture(scope="module")
def workload():
    from dclaw_amd import workloads as W
    return W.fractal_sources(4096, 4096, dx=2.0, seed=1)


def test_fractal4096_first_steps_equal_oracle(dev, oracle_mod, workload):
    from dclaw_amd import step as S
    O = oracle_mod
    pt = workload
    nsteps = 2
    print("[fullsize] oracle 2 steps of 4096x4096", flush=True)
    qo, _, dto, cflo, _ = O.run(O.params_from(pt.params), pt.q, pt.aux, pt.dx, pt.dy, nsteps,
                                1e-3, 0.9, 1e9)
    print("[fullsize] gpu", flush=True)
    r = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy)
    hist = dev.zeros(nsteps, dtype=dev.float64, device="cuda")
    r.run(nsteps, 1e-3, 0.9, 1e9, cfl_hist=hist)
    dev.cuda.synchronize()
    qg = r.q()
    I = (slice(None), slice(2, -2), slice(2, -2))
    assert mask_equal(qg[I], qo[I], pt.params.dry_tolerance)
    err = rel_linf_wet(qg[I], qo[I], pt.params.dry_tolerance)
    print("[fullsize] rel err", err, flush=True)
    assert err <= TOL_STEP
    assert np.max(np.abs(hist.cpu().numpy() - cflo) / cflo) <= TOL_STEP
    assert r.dt.item() == pytest.approx(dto, rel=TOL_STEP)


def test_fractal4096_tiling_independent(dev, workload):
    from dclaw_amd import _lib
    from dclaw_amd import step as S
    pt = workload
    L = _lib.lib()
    out = []
    for seg, mode in ((6, "single"), (24, "ordered"), (40, "dry")):
        print(f"[fullsize] 30 steps, segment {seg}, {mode}", flush=True)
        prev = L.dclaw_set_seg_rows(seg)
        try:
            r = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy, split_mode=mode)
            r.run(30, 1e-3, 0.9, 1e9)
            dev.cuda.synchronize()
            out.append((r.q(), r.dt.item()))
        finally:
            L.dclaw_set_seg_rows(prev)
    for q, dt in out[1:]:
        assert np.array_equal(q, out[0][0]) and dt == out[0][1]


def test_lake_at_rest_4096(dev, workload):
    from dclaw_amd import step as S
    from dclaw_amd import workloads as W
    from dclaw_amd.params import to_c
    pt = workload
    b = np.asfortranarray(pt.aux)[0, 2:-2, 2:-2]
    eta = float(b.max()) + 1.0
    h = eta - b
    c = to_c(pt.params)
    rest = W.make_patch(b, h, 0.6, 1.0, pt.params, pt.dx, pt.dy)
    # hydrostatic pore pressure pb = rho_f g h: no excess pressure to relax
    assert np.allclose(rest.q[4][2:-2, 2:-2], c.rho_f * c.grav * h, rtol=1e-15)
    r = S.PatchRunner(rest.q, rest.aux, rest.params, rest.dx, rest.dy)
    r.run(10, 1e-3, 0.9, 1e9)
    dev.cuda.synchronize()
    q = r.q()
    I = (slice(None), slice(2, -2), slice(2, -2))
    d = q[I] - rest.q[I]
    print("[fullsize] lake at rest max |dh|", np.abs(d[0]).max(), "max |dhu|",
          np.abs(d[1]).max(), flush=True)
    for m in range(5):
        scale = max(np.abs(rest.q[I][m]).max(), 1.0)
        assert np.abs(d[m]).max() <= 1e-11 * scale, m
