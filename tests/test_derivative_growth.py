"""Where the derivative of a long flume run grows (CPU, on the oracle).

A perturbation of the state is pushed through the linearised step (central
differences of the whole step, b4step2 + step2 + update + src2, at a 1e-7
step, renormalised every step) over a flume release of 2000 x 12 cells
(dx = 0.05 m, 31 deg then 2 deg), from step 100 to step 260.  With the
default second-order scheme (MC limiter on all five waves) its size grows by
many orders of magnitude once the flow is supercritical; with the same run
at first order (Godunov, no limited correction) it does not grow.  So the
growth of the run's gradients (the tangent-linear and adjoint steps are the
exact derivative of the step) is a property of the MC-limited second-order
correction's derivative in fast flow: the limited wave phi(r) W with
r = W . W_upwind / |W|^2 has Jacobian entries ~ |W_upwind| / |W| where a
wave family is weak (flux2fw.f:188, limiter; flux2fw.f:215-221, correction).
Neither src2 (its d pb / d pb stays <= 1 on these states) nor the
first-order part expands.
"""
import math

import numpy as np
import pytest

from dclaw_amd import workloads as W

SC = [0.75, 3.0, 3.0, 0.5, 1e4]     # component scales (h, hu, hv, hm, pb)


def _growth(O, order, nsteps=260, start=100):
    pt, _ = W.flume_release(mx=2000, my=12, dx=0.05, seed=3)
    pt.params.order = order
    bc = (1, 1, 1, 1)

    def step(q, aux, dt):
        P = O.params_from(pt.params)
        qb, ab = O.b4step2(P, O.bc_fill(q, bc), aux, pt.dx, pt.dy)
        cfl, fm, fp, gm, gp = O.step2(O.params_from(pt.params), qb, ab, pt.dx, pt.dy, dt)
        out = qb.copy()
        I = slice(2, -2)
        out[:, I, I] = (qb[:, I, I] - dt / pt.dx * (fm[:, 3:-1, I] - fp[:, I, I])
                        - dt / pt.dy * (gm[:, I, 3:-1] - gp[:, I, I]))
        return O.src2(O.params_from(pt.params), out, ab, dt), ab, cfl

    def nrm(v):
        return max(np.abs(v[m]).max() / SC[m] for m in range(5))

    q, aux, dt = O.bc_fill(pt.q, bc), pt.aux, 1e-3
    rng = np.random.default_rng(0)
    d, log10g = None, 0.0
    for n in range(nsteps):
        if n == start:
            d = np.zeros_like(q)
            w = q[0, 2:-2, 2:-2] > 1e-3
            for m in range(5):
                d[m, 2:-2, 2:-2] = rng.normal(size=w.shape) * w * SC[m]
            d /= nrm(d)
        qn, an, cfl = step(q, aux, dt)
        if d is not None:
            a, _, _ = step(q + 1e-7 * d, aux, dt)
            b, _, _ = step(q - 1e-7 * d, aux, dt)
            dd = (a - b) / 2e-7
            log10g += math.log10(nrm(dd))
            d = dd / nrm(dd)
        q, aux, dt = O.bc_fill(qn, bc), an, min(dt * 0.9 / cfl, 1e9)
    return log10g


@pytest.mark.parametrize("order,lo,hi", [(2, 8.0, None), (1, None, 2.0)])
def test_linearised_growth_comes_from_the_limited_correction(oracle_mod, order, lo, hi):
    g = _growth(oracle_mod, order)
    if lo is not None:
        assert g > lo, g
    if hi is not None:
        assert g < hi, g
