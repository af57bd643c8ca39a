"""The bowl-slosh workload's initial state is the reference case's qinit
(tests/bowl-slosh/qinit.f90:75-87 and its init_ptype 0 pressure): Thacker's
solution at t = 0 with cells at or below the dry tolerance zeroed."""
import math

import numpy as np

from dclaw_amd import workloads as W
from helpers import thacker_bowl


def test_bowl_initial_state_is_thacker_at_t0():
    n = 41
    pt = W.bowl_slosh(n)
    q = pt.q[:, 2:-2, 2:-2]
    he, ue, ve, b = thacker_bowl(n, 0.0)
    p = pt.params
    wet = he > p.dry_tolerance
    assert np.array_equal(q[0], np.where(wet, he, 0.0))
    assert ue == 0.0 and not q[1].any()
    omega = math.sqrt(2 * p.grav * W.BOWL_H0) / W.BOWL_A
    assert ve == W.BOWL_SIGMA * omega
    assert np.array_equal(q[2], ve * q[0])
    assert not q[3].any() and not q[5].any() and not q[6].any()
    assert np.array_equal(q[4], p.rho_f * p.grav * q[0])
    assert np.array_equal(pt.aux[0, 2:-2, 2:-2], b)
    # eta > B inside the disc (x - 1/2)^2 + y^2 < 1 - dx^2/6 (a = 1, sigma = 1/2)
    x = -2.0 + (np.arange(n) + 0.5) * 4.0 / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    assert np.all(((X - 0.5) ** 2 + Y ** 2)[wet] < 1.0)
    assert wet.sum() > 0.15 * n * n
    assert p.mref == 0.0 and p.phi_deg == 32.0 and p.dry_tolerance == 1e-3
