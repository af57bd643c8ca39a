"""Synthetic patch workloads (q, aux) in the reference layout.

Layout (the Clawpack contract, kept as is): ``q`` is Fortran-ordered
``q(7, 1-mbc:mx+mbc, 1-mbc:my+mbc)`` with the component fastest; as a numpy
array it has shape ``(7, mx+2mbc, my+2mbc)`` and ``order="F"``.  The same
bytes viewed C-ordered are ``(my+2mbc, mx+2mbc, 7)``, which is how torch
tensors hold it.  ``aux`` is ``aux(maux, ...)`` likewise, with aux(1) = b,
aux(2) = phi, aux(3) = theta and aux(4..6) = fsphi, taudir_x, taudir_y.

Initialisation follows the reference's qinit with init_ptype 3
(qinit.f90:207-283): hm = m0*h, pb = init_pratio*rho_f*g*h, cells with
h <= dry_tolerance zeroed, then qfix (applied here in closed form: pb is
clamped to [0, rho*g*h]).  Ghost bed values copy the nearest interior cell
(setaux.f90:154-184); aux(4..6) start at 0 (setaux.f90:196-198).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List

import numpy as np

from .params import DigParams, config0_params

MEQN = 7
MAUX = 6
MBC = 2


@dataclasses.dataclass
class Patch:
    q: np.ndarray        # (7, mx+2mbc, my+2mbc), Fortran order
    aux: np.ndarray      # (maux, mx+2mbc, my+2mbc), Fortran order
    params: DigParams
    dx: float
    dy: float
    mbc: int = MBC

    @property
    def mx(self) -> int:
        return self.q.shape[1] - 2 * self.mbc

    @property
    def my(self) -> int:
        return self.q.shape[2] - 2 * self.mbc


def _clamp_ghosts(a2: np.ndarray, mbc: int) -> np.ndarray:
    """a2: interior field (mx, my) -> (mx+2mbc, my+2mbc) with ghosts copying
    the nearest interior cell."""
    return np.pad(a2, mbc, mode="edge")


def make_patch(b_int: np.ndarray, h_int: np.ndarray, m_int, pratio_int, params: DigParams,
               dx: float, dy: float, mbc: int = MBC, maux: int = MAUX,
               h_e=None) -> Patch:
    """Build (q, aux) from interior fields of shape (mx, my) (x index first)."""
    mx, my = h_int.shape
    nx, ny = mx + 2 * mbc, my + 2 * mbc
    g = params.grav
    h = np.zeros((nx, ny))
    h[mbc:mbc + mx, mbc:mbc + my] = h_int
    m = np.zeros((nx, ny))
    m[mbc:mbc + mx, mbc:mbc + my] = m_int
    pr = np.zeros((nx, ny))
    pr[mbc:mbc + mx, mbc:mbc + my] = pratio_int
    q = np.zeros((MEQN, nx, ny), order="F")
    hm = m * h
    pb = pr * params.rho_f * g * h
    dry = h <= params.dry_tolerance
    h = np.where(dry, 0.0, h)
    hm = np.where(dry, 0.0, hm)
    pb = np.where(dry, 0.0, pb)
    # qfix's pressure bounds: 0 <= pb <= rho*g*h with rho from m = hm/h
    mm = np.divide(hm, h, out=np.zeros_like(h), where=h > 0)
    rho = params.rho_s * mm + (1.0 - mm) * params.rho_f
    pb = np.minimum(rho * g * h, np.maximum(0.0, pb))
    q[0] = h
    q[3] = hm
    q[4] = pb
    aux = np.zeros((maux, nx, ny), order="F")
    aux[0] = _clamp_ghosts(b_int, mbc)
    aux[1] = params.phi
    aux[2] = 0.0
    if h_e is not None:
        # aux(7) = h_e, the erodible layer depth (digclaw_module.f90:110, i_ent)
        if maux < 7:
            raise ValueError("h_e needs maux >= 7")
        aux[6] = _clamp_ghosts(np.broadcast_to(h_e, (mx, my)).astype(np.float64), mbc)
    return Patch(q=q, aux=aux, params=params, dx=dx, dy=dy, mbc=mbc)


def inclined_plane_mound(n: int = 200, dx: float = 1.0) -> Patch:
    """200x200 cells, dx = 1 m: plane inclined 31 deg in x for the first 120 m
    then a flat runout; Gaussian mound of peak 5 m, sigma 12 cells at
    (40 m, 100 m); u = v = 0; m = 0.62; pb at 0.8 of hydrostatic."""
    p = config0_params()
    x = (np.arange(n) + 0.5) * dx
    y = (np.arange(n) + 0.5) * dx
    X, Y = np.meshgrid(x, y, indexing="ij")
    b = np.where(X < 120.0, math.tan(math.radians(31.0)) * (120.0 - X), 0.0)
    sig = 12.0 * dx
    h = 5.0 * np.exp(-((X - 40.0) ** 2 + (Y - 100.0) ** 2) / (2.0 * sig ** 2))
    return make_patch(b, h, 0.62, 0.8, p, dx, dx)


BOWL_A, BOWL_SIGMA, BOWL_H0 = 1.0, 0.5, 0.1


def bowl_slosh(n: int = 41, dry_tolerance: float = 1.0e-3) -> Patch:
    """Clear water sloshing in a parabolic bowl (Thacker's planar surface),
    the reference's tests/bowl-slosh case: domain [-2, 2]^2 in n x n cells
    (setrun.py:70-79), bed B = h0((x^2 + y^2)/a^2 - 1) with a = 1, h0 = 0.1
    (maketopo.py), surface eta = (sigma h0/a^2)(2x - sigma) with sigma = 0.5,
    h = max(0, eta - B), hu = 0, hv = sigma omega h, omega = sqrt(2 g h0)/a
    (qinit.f90:75-87), m = 0 and pb hydrostatic (init_ptype 0,
    qinit.f90:236-247), and the D-Claw parameters of setrun.py:346-396
    (phi 32 deg, mu 0.005, m0 = 0, kref 1e-8, alpha_c 0.05, delta 0.001).
    The bed is the cell average of the paraboloid, which the reference's
    topo integral over its 200 x 200 topo file approximates; depth and
    momentum are set at cell centres, as qinit.f90 sets them."""
    p = DigParams(grav=9.81, dry_tolerance=dry_tolerance, rho_f=1000.0, rho_s=2700.0,
                  phi_deg=32.0, mu=0.005, mref=0.0, m_crit=0.64, kref=1.0e-8, alpha_c=0.05,
                  delta=0.001, sigma_0=1.0e3, c1=1.0)
    a, sig, h0 = BOWL_A, BOWL_SIGMA, BOWL_H0
    dx = 4.0 / n
    x = -2.0 + (np.arange(n) + 0.5) * dx
    X, Y = np.meshgrid(x, x, indexing="ij")
    b = h0 * ((X * X + Y * Y + dx * dx / 6.0) / (a * a) - 1.0)
    h = np.maximum(0.0, sig * h0 / (a * a) * (2.0 * X - sig) - b)
    pt = make_patch(b, h, 0.0, 1.0, p, dx, dx)
    pt.q[2] = sig * math.sqrt(2.0 * p.grav * h0) / a * pt.q[0]
    return pt


def wet_plane(n: int = 2048, dx: float = 2.0, slope_deg: float = 20.0, h: float = 2.0) -> Patch:
    """Fully wet n x n patch: a plane inclined slope_deg in x under a uniform
    layer of depth h (m = 0.6, pb at 0.8 of hydrostatic), starting at rest.
    Every cell is wet, so every edge runs the full Riemann solver: the
    FP64-bound regime of the step."""
    p = config0_params()
    x = (np.arange(n) + 0.5) * dx
    b = -math.tan(math.radians(slope_deg)) * x[:, None] + np.zeros((n, n))
    return make_patch(b, np.full((n, n), h), 0.6, 0.8, p, dx, dx)


def flume_release(mx: int = 2000, my: int = 40, dx: float = 0.05, seed: int = 3,
                  ngauges: int = 8):
    """Gradient workload: a 2000 x 40 flume of dx = 0.05 m, bed slope 31 deg
    over the first 1600 cells then 2 deg (continuous), a block of 6 m^3 in
    cells 0-79 x 0-39 (h = 6 / (80 * 40 * dx^2) = 0.75 m, m0 = 0.6, pb at 0.8
    of hydrostatic; the pore-pressure ratio is this workload's choice)
    released at t = 0.  Other sizes keep the proportions (break at 80 % of
    the length, block over the first 4 %, depth 0.75 m).  Returns (patch,
    gauge (x, y) positions seeded along the centre line over 30-95 % of the
    length)."""
    rng = np.random.default_rng(seed)
    p = config0_params()
    x = (np.arange(mx) + 0.5) * dx
    xb = int(0.8 * mx) * dx
    b1 = math.tan(math.radians(2.0)) * (mx * dx - x) + np.where(
        x < xb, math.tan(math.radians(31.0)) * (xb - x), 0.0)
    b = b1[:, None] + np.zeros((mx, my))
    h = np.zeros((mx, my))
    h[:max(1, mx // 25), :] = 0.75
    pt = make_patch(b, h, 0.6, 0.8, p, dx, dx)
    gx = np.sort(rng.uniform(0.3 * mx * dx, 0.95 * mx * dx, ngauges))
    gy = np.full(ngauges, 0.5 * my * dx)
    return pt, np.stack([gx, gy], axis=1)


def fractal_surface(mx: int, my: int, rng: np.random.Generator, relief: float = 200.0,
                    slope: float = -2.0) -> np.ndarray:
    """Isotropic random surface with power spectrum ~ k^slope, rescaled to the
    given max-min relief."""
    kx = np.fft.fftfreq(mx)[:, None]
    ky = np.fft.rfftfreq(my)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    amp = k ** (slope / 2.0)
    amp[0, 0] = 0.0
    phase = rng.uniform(0.0, 2.0 * np.pi, size=amp.shape)
    spec = amp * np.exp(1j * phase)
    f = np.fft.irfft2(spec, s=(mx, my))
    f -= f.min()
    rng_ = f.max()
    if rng_ > 0:
        f *= relief / rng_
    return f


def fractal_sources(mx: int = 4096, my: int = 4096, dx: float = 2.0, seed: int = 1,
                    nsrc: int = 64) -> Patch:
    """Single large patch: seeded fractal bed (spectral slope -2, 200 m relief)
    on a 20 deg regional tilt in x, with `nsrc` seeded elliptical masses,
    h ~ U[1,8] m, m ~ U[0.55,0.64], pb/(rho_f g h) ~ U[0.5,1.0]."""
    rng = np.random.default_rng(seed)
    p = config0_params()
    x = (np.arange(mx) + 0.5) * dx
    b = fractal_surface(mx, my, rng) - math.tan(math.radians(20.0)) * x[:, None]
    h = np.zeros((mx, my))
    m = np.full((mx, my), 0.6)
    pr = np.full((mx, my), 0.8)
    ii = np.arange(mx)[:, None]
    jj = np.arange(my)[None, :]
    for _ in range(nsrc):
        cx, cy = rng.uniform(0.05 * mx, 0.95 * mx), rng.uniform(0.05 * my, 0.95 * my)
        a = rng.uniform(10.0, 60.0) * mx / 4096.0 + 2.0
        bb = rng.uniform(10.0, 60.0) * my / 4096.0 + 2.0
        ang = rng.uniform(0.0, np.pi)
        hh, mm, pp = rng.uniform(1.0, 8.0), rng.uniform(0.55, 0.64), rng.uniform(0.5, 1.0)
        i0, i1 = int(max(0, cx - max(a, bb) - 1)), int(min(mx, cx + max(a, bb) + 2))
        j0, j1 = int(max(0, cy - max(a, bb) - 1)), int(min(my, cy + max(a, bb) + 2))
        di = ii[i0:i1] - cx
        dj = jj[:, j0:j1] - cy
        ca, sa = math.cos(ang), math.sin(ang)
        r2 = ((di * ca + dj * sa) / a) ** 2 + ((-di * sa + dj * ca) / bb) ** 2
        inside = r2 <= 1.0
        sub = h[i0:i1, j0:j1]
        sub[inside] = hh
        m[i0:i1, j0:j1][inside] = mm
        pr[i0:i1, j0:j1][inside] = pp
    return make_patch(b, h, m, pr, p, dx, dx)


@dataclasses.dataclass
class Member:
    patch: Patch
    seed: int


def ensemble(n_members: int = 256, lo: int = 64, hi: int = 512, seed: int = 7,
             dx: float = 2.0) -> List[Member]:
    """Particle-filter style ensemble: sides log-uniform in [lo, hi], each
    member with its own seeded fractal bed and one Gaussian source; per-member
    phi ~ U[32,42] deg, m0 ~ U[0.55,0.64], log10 kref ~ U[-11,-9]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_members):
        mx = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        my = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        mrng = np.random.default_rng([seed, k])
        p = config0_params()
        p.phi_deg = float(mrng.uniform(32.0, 42.0))
        m0 = float(mrng.uniform(0.55, 0.64))
        p.kref = float(10.0 ** mrng.uniform(-11.0, -9.0))
        x = (np.arange(mx) + 0.5) * dx
        b = fractal_surface(mx, my, mrng, relief=0.2 * max(mx, my) * dx / 4.0) \
            - math.tan(math.radians(20.0)) * x[:, None]
        X, Y = np.meshgrid(x, (np.arange(my) + 0.5) * dx, indexing="ij")
        cx = mrng.uniform(0.2, 0.5) * mx * dx
        cy = mrng.uniform(0.3, 0.7) * my * dx
        sig = mrng.uniform(0.04, 0.08) * min(mx, my) * dx
        h = mrng.uniform(2.0, 6.0) * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2.0 * sig ** 2))
        out.append(Member(make_patch(b, h, m0, 0.8, p, dx, dx), seed=k))
    return out
