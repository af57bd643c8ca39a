"""Pins of the oracle's time-loop pieces: the dt rule with the output-time hit
(AMRClaw tick.f:135-144, :499) and the gauge interpolation (the flume port's
convention, gradient_particle_filter/flume_solver.py:91-99, in 2-D), against
closed forms and the rules' own case analysis."""
import numpy as np
import pytest

from dclaw_amd import workloads as W


def test_dt_rule_cases(oracle_mod):
    O = oracle_mod
    # far from the output time: CFL rule and dt_max cap
    t, dn, hit = O.dt_rule(0.01, 0.5, 1.0, 0.9, 1.0, tout=5.0)
    assert (t, dn, hit) == (1.0 + 0.01, 0.01 * 0.9 / 0.5, 0)
    t, dn, hit = O.dt_rule(0.01, 0.001, 1.0, 0.9, 0.05, tout=5.0)
    assert dn == 0.05 and hit == 0
    # cflgrid = 0: dtnew = infinity, capped by dt_max
    t, dn, hit = O.dt_rule(0.01, 0.0, 1.0, 0.9, 0.2, tout=float("inf"))
    assert dn == 0.2 and hit == 0
    # within 1.001 dt of the output time: land on it exactly
    t, dn, hit = O.dt_rule(0.01, 0.5, 1.0, 0.9, 1.0, tout=1.01 + 0.018)
    assert hit == 1 and dn == (1.01 + 0.018) - 1.01
    # a slightly longer step (up to 0.1 %) is accepted to hit the output
    tout = 1.01 + 1.0005 * 0.018
    t, dn, hit = O.dt_rule(0.01, 0.5, 1.0, 0.9, 1.0, tout=tout)
    assert hit == 1 and dn == tout - 1.01 and dn > 0.018
    # beyond 1.001 dt: no hit
    t, dn, hit = O.dt_rule(0.01, 0.5, 1.0, 0.9, 1.0, tout=1.01 + 1.0011 * 0.018)
    assert hit == 0
    # output time already passed: no hit
    t, dn, hit = O.dt_rule(0.01, 0.5, 1.0, 0.9, 1.0, tout=1.0)
    assert hit == 0


def _field(mx, my, f):
    q = np.zeros((7, mx + 4, my + 4), order="F")
    aux = np.zeros((1, mx + 4, my + 4), order="F")
    x = (np.arange(-1, mx + 3) - 0.5)[:, None]
    y = (np.arange(-1, my + 3) - 0.5)[None, :]
    for m in range(7):
        q[m] = f(m, x, y)
    aux[0] = 100.0 + 0.3 * x - 0.2 * y
    return q, aux


def test_gauges_bilinear_is_exact_on_linear_fields(oracle_mod):
    O = oracle_mod
    mx, my = 20, 16
    lin = lambda m, x, y: (5.0 + m) + 0.25 * (m + 1) * x - 0.125 * y   # h > 0 everywhere
    q, aux = _field(mx, my, lin)
    rng = np.random.default_rng(0)
    xg = rng.uniform(0.5, mx - 0.5, 30)
    yg = rng.uniform(0.5, my - 0.5, 30)
    out = O.gauges(q, aux, 0.0, 0.0, 1.0, 1.0, 1e-3, xg, yg)
    for m in range(7):
        np.testing.assert_allclose(out[:, m], lin(m, xg, yg), rtol=1e-13, atol=1e-13)
    eta = lin(0, xg, yg) + (100.0 + 0.3 * xg - 0.2 * yg) - lin(6, xg, yg)
    np.testing.assert_allclose(out[:, 7], eta, rtol=1e-13)


def test_gauges_dry_neighbour_takes_containing_cell(oracle_mod):
    O = oracle_mod
    mx, my = 10, 10
    q, aux = _field(mx, my, lambda m, x, y: 2.0 + 0.1 * m + 0.01 * x + 0.02 * y)
    q[0, 2 + 4, 2 + 5] = 0.0         # cell (5, 6) dry (0-based with 2 ghosts: i+1 -> index)
    # gauge at (4.7, 5.2): lower-left centre cell (5, 5), neighbours (6, 5), (5, 6), (6, 6);
    # containing cell (5, 6)
    out = O.gauges(q, aux, 0.0, 0.0, 1.0, 1.0, 1e-3, [4.7], [5.2])
    ic, jc = 5, 6
    assert np.array_equal(out[0, :7], q[:, ic + 1, jc + 1])


def test_gauges_1d_flume_convention(oracle_mod):
    """Gauge on an x-edge at a row centre: both neighbours wet -> the two-cell
    average; one dry -> the cell containing the gauge (the right one)."""
    O = oracle_mod
    mx, my = 12, 3
    rng = np.random.default_rng(4)
    q = np.asfortranarray(rng.uniform(1.0, 2.0, (7, mx + 4, my + 4)))
    aux = np.zeros((1, mx + 4, my + 4), order="F")
    xg, yg = [6.0], [1.5]            # edge between cells 6 and 7, centre of row 2
    out = O.gauges(q, aux, 0.0, 0.0, 1.0, 1.0, 1e-3, xg, yg)
    qa, qb = q[:, 6 + 1, 2 + 1], q[:, 7 + 1, 2 + 1]
    assert np.array_equal(out[0, :7], 0.5 * (qa + qb))
    q[0, 6 + 1, 2 + 1] = 0.0
    out = O.gauges(q, aux, 0.0, 0.0, 1.0, 1.0, 1e-3, xg, yg)
    assert np.array_equal(out[0, :7], q[:, 7 + 1, 2 + 1])


def test_run_patches_equals_run_per_patch(oracle_mod):
    """The OpenMP-over-patches loop of the CPU baseline (the "omp" build)
    gives each patch exactly the run it gets alone."""
    from dclaw_amd import workloads as W
    O = oracle_mod
    ms = W.ensemble(6, 24, 60, seed=3)
    P = [O.params_from(m.patch.params) for m in ms]
    with O.libm("omp"):
        qs, dts = O.run_patches(P, [m.patch.q for m in ms], [m.patch.aux for m in ms],
                                [m.patch.dx for m in ms], [m.patch.dy for m in ms], 4, 1e-3,
                                0.9, 1e9, nthreads=4)
    for k, m in enumerate(ms):
        q, _, dt, _, _ = O.run(O.params_from(m.patch.params), m.patch.q, m.patch.aux,
                               m.patch.dx, m.patch.dy, 4, 1e-3, 0.9, 1e9)
        assert np.array_equal(q, qs[k]) and dt == dts[k], k
