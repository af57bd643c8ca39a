"""Test helpers: small synthetic patches, the lane-emulator binding and the
parity metrics used by every comparison against the oracle."""
from __future__ import annotations

import ctypes
import math

import numpy as np

from dclaw_amd import workloads as W
from dclaw_amd.params import CParams, to_c

_dp = ctypes.POINTER(ctypes.c_double)

# Parity bars: per step 1e-12, after 1000 steps 1e-9, relative
# L-inf over wet cells (normwise per component); wet/dry mask bit-identical.
TOL_STEP = 1e-12
TOL_LONG = 1e-9


def slope_patch(mx=40, my=30, slope=0.3, phi_deg=40.0, seed=0, h0=3.0, m0=0.62, pratio=0.8):
    """A mound on a tilted, rippled bed; a steep slope and low phi make it flow."""
    rng = np.random.default_rng(seed)
    p = W.config0_params()
    p.phi_deg = phi_deg
    x = np.arange(mx) + 0.5
    y = np.arange(my) + 0.5
    X, Y = np.meshgrid(x, y, indexing="ij")
    b = slope * (mx - X) + 0.1 * np.sin(0.7 * Y) + 0.05 * rng.standard_normal((mx, my))
    cx, cy = mx * 0.35, my * 0.45
    sig = max(3.0, min(mx, my) / 6.0)
    h = h0 * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * sig ** 2))
    return W.make_patch(b, h, m0, pratio, p, 1.0, 1.0)


def thacker_bowl(n: int, t: float):
    """Thacker's exact planar-surface solution in the bowl of
    workloads.bowl_slosh at time t, at the n x n cell centres of [-2, 2]^2
    (docs/src/tests.rst:64-84): (h, u, v, bed)."""
    a, sig, h0, g = W.BOWL_A, W.BOWL_SIGMA, W.BOWL_H0, 9.81
    om = math.sqrt(2.0 * g * h0) / a
    dx = 4.0 / n
    x = -2.0 + (np.arange(n) + 0.5) * dx
    X, Y = np.meshgrid(x, x, indexing="ij")
    b = h0 * ((X * X + Y * Y + dx * dx / 6.0) / (a * a) - 1.0)
    eta = sig * h0 / (a * a) * (2.0 * X * math.cos(om * t) + 2.0 * Y * math.sin(om * t) - sig)
    return (np.maximum(0.0, eta - b), -sig * om * math.sin(om * t), sig * om * math.cos(om * t),
            b)


def bowl_errors(q: np.ndarray, n: int, t: float, deep: float = 0.02):
    """Distance of an interior bowl state q (7, n, n) from Thacker's solution
    at time t: mean |h - h_exact| over the domain and the largest velocity
    error over the cells deeper than `deep` in both."""
    he, ue, ve, _ = thacker_bowl(n, t)
    h = q[0]
    d = (he > deep) & (h > deep)
    u = q[1][d] / h[d]
    v = q[2][d] / h[d]
    return float(np.abs(h - he).mean()), float(max(np.abs(u - ue).max(), np.abs(v - ve).max()))


def rel_linf_wet(a: np.ndarray, b: np.ndarray, drytol: float) -> float:
    """max over components of ||a-b||_inf / ||b||_inf over cells wet in b."""
    wet = b[0] > drytol
    if not wet.any():
        return float(np.abs(a - b).max())
    worst = 0.0
    for m in range(a.shape[0]):
        bb = b[m][wet]
        scale = np.abs(bb).max()
        d = np.abs(a[m][wet] - bb).max()
        if scale == 0.0:
            worst = max(worst, d)
        else:
            worst = max(worst, d / scale)
    return worst


def mask_equal(a: np.ndarray, b: np.ndarray, drytol: float) -> bool:
    return bool(np.array_equal(a[0] > drytol, b[0] > drytol))


class Emu:
    """The product kernel body run by the host lane emulator."""

    def __init__(self, lib):
        self.L = lib
        lib.emu_stepgrid.argtypes = [ctypes.POINTER(CParams), _dp, _dp, _dp, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int), _dp]

    def stepgrid(self, params, q, aux, dx, dy, dt, seg_rows=32, write_aux=True, bc=None):
        cp = to_c(params)
        q = np.array(q, dtype=np.float64, order="F", copy=True)
        qo = np.zeros_like(q, order="F")
        aux = np.array(aux, dtype=np.float64, order="F", copy=True)
        cfl = ctypes.c_double(0.0)
        mx, my = q.shape[1] - 4, q.shape[2] - 4
        self.L.emu_stepgrid(ctypes.byref(cp), q.ctypes.data_as(_dp), qo.ctypes.data_as(_dp),
                            aux.ctypes.data_as(_dp), aux.shape[0], int(write_aux), mx, my, dx,
                            dy, dt, seg_rows,
                            None if bc is None else (ctypes.c_int * 4)(*bc), ctypes.byref(cfl))
        return qo, aux, cfl.value


def advance_oracle(O, patch, nsteps, dt0=1e-3, cfl=0.9, dt_max=1e9, bc=(1, 1, 1, 1)):
    """Run the oracle a few steps to get a moving state; returns (q with
    ghosts filled, aux, next dt)."""
    P = O.params_from(patch.params)
    q = O.bc_fill(patch.q, bc)
    q, aux, dt, _, _ = O.run(P, q, patch.aux, patch.dx, patch.dy, nsteps, dt0, cfl, dt_max, bc)
    return O.bc_fill(q, bc), aux, dt
