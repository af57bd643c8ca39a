"""GPU: the device-resident time loop pieces (dt rule with output times,
gauges, check4nans in the loop) against the oracle's restatements, and a
run to output times with gauges against the same loop on the oracle."""
import numpy as np
import pytest

from helpers import advance_oracle, mask_equal, slope_patch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dclaw_amd import _lib
    _lib.lib()
    return torch


def test_dt_update_out_equals_oracle(dev, oracle_mod):
    from dclaw_amd import step as S
    O = oracle_mod
    rng = np.random.default_rng(2)
    n = 500
    dt = rng.uniform(1e-4, 1e-2, n)
    cfl = rng.uniform(0.0, 1.2, n)
    cfl[::7] = 0.0
    t = rng.uniform(0.0, 1.0, n)
    tnext = t + dt
    tout = tnext + rng.uniform(0.0, 0.05, n)
    tout[::5] = np.inf
    tout[::11] = tnext[::11] - 1e-3          # already passed
    cd, dmax = 0.9, 0.02
    dd, cc, tt = (dev.tensor(a, dtype=dev.float64, device="cuda") for a in (dt, cfl, t))
    to = dev.tensor(tout, dtype=dev.float64, device="cuda")
    hit = dev.zeros(n, dtype=dev.int32, device="cuda")
    S.dt_update_out(dd, cc, tt, to, hit, cd, dmax)
    ref = [O.dt_rule(dt[k], cfl[k], t[k], cd, dmax, tout[k]) for k in range(n)]
    assert np.array_equal(tt.cpu().numpy(), [r[0] for r in ref])
    assert np.array_equal(dd.cpu().numpy(), [r[1] for r in ref])
    assert np.array_equal(hit.cpu().numpy(), [r[2] for r in ref])
    assert 0 < sum(r[2] for r in ref) < n
    assert not cc.any()


def test_gauges_equal_oracle(dev, oracle_mod):
    from dclaw_amd import step as S
    O = oracle_mod
    pt = slope_patch(80, 50, slope=0.8, phi_deg=25.0, seed=6)
    q, aux, _ = advance_oracle(O, pt, 8)
    rng = np.random.default_rng(1)
    xy = np.stack([rng.uniform(0.5, 79.5, 40), rng.uniform(0.5, 49.5, 40)], axis=1)
    xy[:5] = [[10.0, 10.0], [10.5, 10.5], [30.0, 20.5], [0.5, 0.5], [79.49, 49.49]]
    qd = S.to_device_q(q)
    out = dev.zeros((40, 8), dtype=dev.float64, device="cuda")
    S.gauges(qd, S.bed_tensor(aux), 1, 80, 50, 0.0, 0.0, 1.0, 1.0, pt.params.dry_tolerance,
             dev.tensor(xy, dtype=dev.float64, device="cuda"), out)
    ref = O.gauges(q, aux, 0.0, 0.0, 1.0, 1.0, pt.params.dry_tolerance, xy[:, 0], xy[:, 1])
    assert np.array_equal(out.cpu().numpy(), ref)
    wet = ref[:, 0] > pt.params.dry_tolerance
    assert wet.any() and (~wet).any()


def _oracle_run_until(O, pt, outs, dt0, cd, dmax, xy, bc=(1, 1, 1, 1)):
    """The same loop on the oracle, in AMRClaw tick's order."""
    P = O.params_from(pt.params)
    q, aux, t = pt.q, pt.aux, 0.0
    target, landing = 0, None
    if 1.001 * dt0 >= outs[0]:
        dt0, landing, target = outs[0], 0, 1
    dt, reached, times, gv = dt0, [], [], []
    while True:
        q = O.bc_fill(q, bc)
        q, aux, cfl = O.stepgrid(P, q, aux, pt.dx, pt.dy, dt)
        tout = outs[target] if target < len(outs) else np.inf
        t, dt, hit = O.dt_rule(dt, cfl, t, cd, dmax, tout)
        times.append(t)
        gv.append(O.gauges(q, aux, 0.0, 0.0, pt.dx, pt.dy, pt.params.dry_tolerance, xy[:, 0],
                           xy[:, 1]))
        if landing is not None:
            reached.append(t)
            if landing == len(outs) - 1:
                break
            landing = None
        if hit:
            landing, target = target, target + 1
    return q, reached, np.array(times), np.array(gv)


def test_run_until_outputs_and_gauges_equal_oracle(dev, oracle_mod):
    from dclaw_amd import step as S
    O = oracle_mod
    pt = slope_patch(64, 40, slope=0.9, phi_deg=22.0, seed=3)
    outs = [0.02, 0.05, 0.0501, 0.12]
    xy = np.array([[20.0, 18.5], [25.25, 20.75], [40.0, 30.0]])
    print("[run_until] 64x40 outputs", outs, flush=True)
    r = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy)
    seen = []
    res = r.run_until(outs, 1e-3, 0.9, 1e9, gauges_xy=xy, check_nans=True,
                      on_output=lambda run, tout: seen.append(tout))
    qo, reached, times, gv = _oracle_run_until(O, pt, outs, 1e-3, 0.9, 1e9, xy)
    assert seen == outs
    assert res.outputs == reached
    assert np.allclose(res.outputs, outs, rtol=1e-14)
    assert np.array_equal(res.times, times)
    assert np.array_equal(res.gauge_values, gv)
    I = (slice(None), slice(2, -2), slice(2, -2))
    assert np.array_equal(r.q()[I], qo[I])
    assert mask_equal(r.q()[I], qo[I], pt.params.dry_tolerance)


def test_run_until_reports_nan(dev):
    """A NaN placed in the initial state (a dry cell's hchi, which qfix zeroes)
    is reported by check4nans at the first output."""
    from dclaw_amd import _lib
    from dclaw_amd import step as S
    pt = slope_patch(30, 20, seed=2)
    q = np.array(pt.q, order="F", copy=True)
    dry = np.argwhere(q[0, 2:-2, 2:-2] <= pt.params.dry_tolerance)[0] + 2
    q[5, dry[0], dry[1]] = np.nan
    r = S.PatchRunner(q, pt.aux, pt.params, pt.dx, pt.dy)
    with pytest.raises(_lib.DClawError, match="check4nans"):
        r.run_until([0.01], 1e-3, 0.9, 1e9, check_nans=True)


def test_runner_frames_round_trip(dev, oracle_mod, tmp_path):
    """Frames written from the device state at the output times read back as
    the oracle's state (ghosts filled as bound() fills them) and eta."""
    from dclaw_amd import io as DIO
    from dclaw_amd import step as S
    O = oracle_mod
    pt = slope_patch(40, 30, slope=0.9, phi_deg=22.0, seed=8)
    outs = [0.01, 0.03]
    r = S.PatchRunner(pt.q, pt.aux, pt.params, pt.dx, pt.dy)
    frames = []

    def out(run, tout):
        k = len(frames) + 1
        run.write_frame(str(tmp_path), k, tout, "binary64")
        frames.append(k)

    r.run_until(outs, 1e-3, 0.9, 1e9, on_output=out)
    assert frames == [1, 2]
    t, grids = DIO.read_frame(str(tmp_path), 2)
    assert t == 0.03
    qh = O.bc_fill(r.q(), (1, 1, 1, 1))
    eta = qh[0] + pt.aux[0] - qh[6]
    assert np.array_equal(grids[0]["qeta"][:7], qh)
    assert np.array_equal(grids[0]["qeta"][7], eta)
