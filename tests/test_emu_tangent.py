"""The tangent-linear step (dclaw_tangent.h), run on the CPU by the lane
emulator: its values equal the step's bit for bit, and its tangent equals
central finite differences of the step (run by the same emulator) along
parameter and state directions, for one step and for a chain of steps on a
fixed time grid.

Finite differences are taken at a relative step of 1e-5 and compared per
component against that component's largest difference (pore pressure is
~1e4 times the other components, so a common scale would hide it); a
tangent is a point derivative, so a cell whose finite difference straddles
a branch of the scheme (the static-friction test, the limiter) can differ,
and the tolerance allows a few such cells."""
import ctypes

import numpy as np
import pytest

from dclaw_amd.params import CParams, to_c
from helpers import advance_oracle, slope_patch

_dp = ctypes.POINTER(ctypes.c_double)
I = (slice(None), slice(2, -2), slice(2, -2))


def _fwd(lib, c, q, aux, dx, dy, dt, seg, bc):
    qo = np.zeros_like(q, order="F")
    a2 = np.array(aux, order="F", copy=True)
    cfl = ctypes.c_double(0.0)
    mx, my = q.shape[1] - 4, q.shape[2] - 4
    lib.emu_stepgrid(ctypes.byref(c), q.ctypes.data_as(_dp), qo.ctypes.data_as(_dp),
                     a2.ctypes.data_as(_dp), a2.shape[0], 0, mx, my, ctypes.c_double(dx),
                     ctypes.c_double(dy), ctypes.c_double(dt), seg,
                     None if bc is None else (ctypes.c_int * 4)(*bc), ctypes.byref(cfl))
    return qo, cfl.value


def _tl(lib, c, dc, q, dq, aux, dx, dy, dt, seg, bc):
    qo = np.zeros_like(q, order="F")
    dqo = np.zeros_like(q, order="F")
    cfl = ctypes.c_double(0.0)
    mx, my = q.shape[1] - 4, q.shape[2] - 4
    assert lib.emu_stepgrid_tl(ctypes.byref(c), ctypes.byref(dc), q.ctypes.data_as(_dp),
                               dq.ctypes.data_as(_dp), qo.ctypes.data_as(_dp),
                               dqo.ctypes.data_as(_dp), aux.ctypes.data_as(_dp), aux.shape[0], mx,
                               my, ctypes.c_double(dx), ctypes.c_double(dy), ctypes.c_double(dt),
                               seg, None if bc is None else (ctypes.c_int * 4)(*bc),
                               ctypes.byref(cfl)) == 0
    return qo, dqo, cfl.value


def _compare(tl, fd, qref, eps, max_bad=4, tol=1e-4):
    """Per component m: |tl - fd| <= tol * max|fd_m| + the roundoff floor of
    a central difference at step eps (8 ulp of max|q_m| / eps), except in at
    most max_bad cells; returns the worst median relative error."""
    worst, bad = 0.0, 0
    for m in range(7):
        scale = max(np.abs(fd[m]).max(), np.abs(tl[m]).max())
        if scale == 0.0:
            continue
        floor = 8.0 * np.abs(qref[m]).max() * np.finfo(float).eps / eps
        err = np.abs(tl[m] - fd[m])
        bad += int((err > tol * scale + floor).sum())
        worst = max(worst, float(np.median(err / scale)))
    assert bad <= max_bad, f"{bad} cells beyond {tol}"
    return worst


@pytest.fixture(scope="module")
def moving(oracle_mod):
    pt = slope_patch(70, 40, slope=0.9, phi_deg=22.0, seed=11)
    q, aux, dt = advance_oracle(oracle_mod, pt, 8)
    return pt, np.asfortranarray(q), np.asfortranarray(aux), dt


@pytest.mark.parametrize("bc", [None, (3, 3, 3, 3), (1, 3, 1, 3)])
def test_values_are_the_step(emu_lib, moving, bc):
    pt, q, aux, dt = moving
    c = to_c(pt.params)
    dq = np.asfortranarray(np.random.default_rng(1).normal(size=q.shape))
    qt, dqt, cflt = _tl(emu_lib, c, CParams(phi=1.0), q, dq, aux, pt.dx, pt.dy, dt, 8, bc)
    qf, cflf = _fwd(emu_lib, c, q, aux, pt.dx, pt.dy, dt, 8, bc)
    assert np.array_equal(qt, qf) and cflt == cflf
    assert np.abs(dqt[I]).max() > 0.0
    # a zero direction gives a zero tangent
    _, dq0, _ = _tl(emu_lib, c, CParams(), q, np.zeros_like(q, order="F"), aux, pt.dx, pt.dy, dt, 8,
                    bc)
    assert not dq0.any()


@pytest.mark.parametrize("name", ["phi", "kref", "mu", "delta", "alpha_c", "m_crit", "rho_s",
                                  "sigma_0", "c1"])
def test_parameter_tangent_matches_finite_differences(emu_lib, moving, name):
    pt, q, aux, dt = moving
    c = to_c(pt.params)
    base = getattr(c, name)
    eps = 1e-5 * abs(base)
    _, dqt, _ = _tl(emu_lib, c, CParams(**{name: 1.0}), q, np.zeros_like(q, order="F"), aux, pt.dx,
                    pt.dy, dt, 8, None)
    c1, c2 = to_c(pt.params), to_c(pt.params)
    setattr(c1, name, base + eps)
    setattr(c2, name, base - eps)
    fd = (_fwd(emu_lib, c1, q, aux, pt.dx, pt.dy, dt, 8, None)[0]
          - _fwd(emu_lib, c2, q, aux, pt.dx, pt.dy, dt, 8, None)[0]) / (2 * eps)
    assert np.abs(fd[I]).max() > 0.0
    assert _compare(dqt[I], fd[I], q[I], eps) < 1e-6


@pytest.mark.parametrize("variant", [
    dict(segregation=1, delta_kr_order=1.5), dict(src2method=1), dict(alphamethod=1),
    dict(friction_forcing=1, manning_coefficient=0.06), dict(src2method=-1),
])
def test_tangent_through_other_branches(emu_lib, oracle_mod, variant):
    """Segregation (pow), the other pressure-relaxation and alpha closures,
    Manning friction: d/dphi against finite differences."""
    pt = slope_patch(60, 36, slope=0.9, phi_deg=22.0, seed=5)
    for k, v in variant.items():
        setattr(pt.params, k, v)
    if variant.get("segregation"):
        pt.q[5] = 0.4 * pt.q[0]
    q, aux, dt = advance_oracle(oracle_mod, pt, 6)
    q, aux = np.asfortranarray(q), np.asfortranarray(aux)
    c = to_c(pt.params)
    name = "manning_coefficient" if variant.get("friction_forcing") else "phi"
    base = getattr(c, name)
    eps = 1e-5 * abs(base)
    _, dqt, _ = _tl(emu_lib, c, CParams(**{name: 1.0}), q, np.zeros_like(q, order="F"), aux, pt.dx,
                    pt.dy, dt, 8, None)
    c1, c2 = to_c(pt.params), to_c(pt.params)
    setattr(c1, name, base + eps)
    setattr(c2, name, base - eps)
    fd = (_fwd(emu_lib, c1, q, aux, pt.dx, pt.dy, dt, 8, None)[0]
          - _fwd(emu_lib, c2, q, aux, pt.dx, pt.dy, dt, 8, None)[0]) / (2 * eps)
    assert np.abs(fd[I]).max() > 0.0
    assert _compare(dqt[I], fd[I], q[I], eps) < 1e-6


def test_state_tangent_over_a_chain_of_steps(emu_lib, moving):
    """Three steps on a fixed time grid, tangent along a random perturbation of
    h, hu, hv, hm and pb in the wet cells, against central differences of the
    chain.  (hchi and bdif are 0 here, at the bounds qfix and b4step2 clip
    them to, where a central difference straddles the kink.)"""
    pt, q, aux, dt = moving
    c = to_c(pt.params)
    rng = np.random.default_rng(3)
    dq0 = np.zeros_like(q, order="F")
    wet = q[0] > 0.1
    for m in range(5):
        dq0[m][wet] = rng.normal(size=wet.sum()) * (np.abs(q[m][wet]).max() + 1.0)
    dq0 = np.asfortranarray(dq0)
    dts = [dt, 0.9 * dt, 1.1 * dt]
    qt, dqt = q, dq0
    for d in dts:
        qt, dqt, _ = _tl(emu_lib, c, CParams(), qt, dqt, aux, pt.dx, pt.dy, d, 8, (1, 1, 1, 1))
    # a random perturbation of every wet cell meets the limiter's and the
    # friction law's kinks within a few cells' worth of eps; 1e-7 keeps the
    # central difference on one side of them
    eps = 1e-7

    def chain(q0):
        x = np.asfortranarray(q0)
        for d in dts:
            x, _ = _fwd(emu_lib, c, x, aux, pt.dx, pt.dy, d, 8, (1, 1, 1, 1))
        return x
    fd = (chain(q + eps * dq0) - chain(q - eps * dq0)) / (2 * eps)
    assert _compare(dqt[I], fd[I], q[I], eps, max_bad=8) < 1e-5
