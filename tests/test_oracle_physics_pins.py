"""Pins of the oracle's physics that bit parity with the GPU cannot give.

The GPU step is bit-identical to the oracle, so parity proves the GPU's
scheduling and summation order; whether the oracle restates the reference's
physics correctly is pinned here, by consequences of the governing equations
and of the numerical method that a plausible transcription slip breaks:

* the order of accuracy of the wave-propagation scheme on smooth water flow
  (second order with the limited Lax-Wendroff correction, first order
  without it; LeVeque 2002, ch. 6 and 20), in one and two dimensions;
* the eigenvectors of the transverse flux Jacobian built from the governing
  equations in docs/src/theory.md:72-111 (the pore-pressure equation gives
  the pressure component rho g varrho of the two gravity waves, the
  solid-mass equation the component m);
* in the Riemann solver, the wave that travels against the interface flow
  carries exactly its own side's eigenvector ratios (the contact jumps of v,
  m, pb, chi ride with the flow: riemannsolvers_dclaw.f:471-517), and a
  constant-discharge flow over a bed step at the solver's own stationary
  depth jump makes no waves (the bed-source and jump lines,
  riemannsolvers_dclaw.f:197-265, must agree);
* the source rates of docs/src/theory.md:139-208 (relaxation zeta D,
  dilatancy 3 |v| tan(psi) / (alpha h), the solid-mass and depth rates of
  D, the critical-state m_eq) at vanishing time steps (Richardson
  extrapolation over two steps);
* b4step2's speed limit (b4step2.f90:74-81) and the pore-pressure bounds of
  qfix_cmass (digclaw_module.f90:353-387);
* the reference's bowl-slosh case (tests/bowl-slosh): clear water in a
  paraboloid bowl converges to Thacker's exact planar-surface solution
  (curved bed, moving shoreline, rotating flow; docs/src/tests.rst:56-84).

tests/test_oracle_mutants.py compiles altered copies of the oracle (one
plausible slip each) and checks that these pins reject every one of them.
The pins are plain functions of the oracle module so that it can run them.
"""
import math

import numpy as np
import pytest

from dclaw_amd import workloads as W
from dclaw_amd.params import config0_params

G = 9.81


def _params(O, **kw):
    p = config0_params()
    for k, v in kw.items():
        setattr(p, k, v)
    return p, O.params_from(p)


# ------------------------------------------------ order of accuracy (flux2)
def _water_run(O, n, order, twod, transverse=2, L=20.0, T=0.8):
    """A smooth water hump (m = 0, hydrostatic pb) on a flat bed, n cells per
    side, to time T at a fixed Courant number ~0.45."""
    p, prm = _params(O, order=order, transverse=transverse)
    dx = L / n
    x = (np.arange(n) + 0.5) * dx
    if twod:
        X, Y = np.meshgrid(x, x, indexing="ij")
        h = 1.0 + 0.1 * np.exp(-((X - 0.45 * L) ** 2 + (Y - 0.55 * L) ** 2) / (2 * (0.12 * L) ** 2))
    else:
        h = (1.0 + 0.1 * np.exp(-((x - 0.5 * L) / (0.1 * L)) ** 2))[:, None] * np.ones((1, 2))
    pt = W.make_patch(np.zeros(h.shape), h, 0.0, 1.0, p, dx, dx)
    q, aux = pt.q, pt.aux
    nsteps = int(math.ceil(T / (0.45 * dx / math.sqrt(G * 1.1))))
    dt = T / nsteps
    for _ in range(nsteps):
        q = O.bc_fill(q, (1, 1, 1, 1))
        q, aux, _ = O.stepgrid(prm, q, aux, dx, dx, dt)
    return q[:, 2:-2, 2:-2]


def _observed_orders(O, ns, order, twod, transverse=2):
    sols = [_water_run(O, n, order, twod, transverse) for n in ns]

    def coarsen(a):
        if twod:
            return 0.25 * (a[:, 0::2, 0::2] + a[:, 1::2, 0::2] + a[:, 0::2, 1::2]
                           + a[:, 1::2, 1::2])
        return 0.5 * (a[:, 0::2, :] + a[:, 1::2, :])
    errs = []
    for a, b in zip(sols[:-1], sols[1:]):
        errs.append(max(np.abs(a[k] - coarsen(b)[k]).mean() for k in (0, 1, 2)))
    return [math.log2(errs[k] / errs[k + 1]) for k in range(len(errs) - 1)]


def pin_second_order_1d(O):
    """order 2 (MC-limited Lax-Wendroff correction, flux2fw.f:195-226):
    observed L1 self-convergence order about 2; order 1 (Godunov): about 1."""
    o2 = _observed_orders(O, [50, 100, 200, 400], 2, False)
    assert min(o2) > 1.8, o2
    o1 = _observed_orders(O, [50, 100, 200], 1, False)
    assert 0.6 < min(o1) and max(o1) < 1.2, o1


def pin_second_order_2d(O):
    """Radially spreading hump, order 2 with transverse corrections of the
    second-order terms (transverse 2, flux2fw.f:236-282) and of the
    first-order fluctuations only (transverse 1): second order both ways."""
    for tr in (2, 1):
        o = _observed_orders(O, [30, 60, 120], 2, True, transverse=tr)
        assert min(o) > 1.8, (tr, o)


# ------------------------------------------------ transverse eigenvectors (rpt2)
def _transverse_jacobian(h, un, ut, m, chi, rho, varrho):
    """Quasi-linear Jacobian of the transverse flux (direction of ut) of
    (h, h un, h ut, hm, pb, h chi) from docs/src/theory.md:72-111 with
    kappa = 1 and a flat bed: mass flux h ut, momentum fluxes h un ut and
    h ut^2 + g h^2/2, hm ut, h chi ut, and the pb equation's transverse terms
    -rho g varrho ut h_y + rho g varrho (h ut)_y + ut pb_y."""
    c = rho * G * varrho
    return np.array([
        [0.0, 0.0, 1.0, 0.0, 0.0, 0.0],
        [-un * ut, ut, un, 0.0, 0.0, 0.0],
        [G * h - ut * ut, 0.0, 2.0 * ut, 0.0, 0.0, 0.0],
        [-m * ut, 0.0, m, ut, 0.0, 0.0],
        [-c * ut, 0.0, c, 0.0, ut, 0.0],
        [-chi * ut, 0.0, chi, 0.0, 0.0, ut],
    ])


def pin_rpt2_gravity_eigenvectors(O):
    """rpt2 maps each gravity-wave eigenvector r (eigenvalue ut -+ sqrt(g h))
    of the transverse Jacobian to lambda r, every component included (pb:
    rho g varrho, hm: m), and the shear eigenvector e_(h un) to ut e_(h un)."""
    p, prm = _params(O)
    for ixy in (1, 2):
        for h, un, ut, m, chi in [(1.3, 0.7, -0.4, 0.6, 0.3), (0.4, -1.1, 0.9, 0.2, 0.8),
                                  (2.0, 0.0, 3.5, 0.55, 0.0)]:
            rho = m * p.rho_s + (1 - m) * p.rho_f
            varrho = (p.rho_f + 3 * rho) / (4 * rho)
            B = _transverse_jacobian(h, un, ut, m, chi, rho, varrho)
            lam, R = np.linalg.eig(B)
            n = 5
            q1d = np.zeros((7, n + 4), order="F")
            iu, iv = (1, 2) if ixy == 1 else (2, 1)     # normal, transverse momentum
            q1d[0], q1d[iu], q1d[iv], q1d[3] = h, h * un, h * ut, h * m
            q1d[4], q1d[5] = 0.7 * p.rho_f * G * h, h * chi
            aux = np.zeros((6, n + 4), order="F")
            targets = [ut - math.sqrt(G * h), ut + math.sqrt(G * h)]
            for lt in targets:
                k = int(np.argmin(np.abs(lam - lt)))
                assert abs(lam[k] - lt) < 1e-9
                r = np.real(R[:, k]) / np.real(R[0, k])
                vec = np.zeros(7)
                vec[[0, iu, iv, 3, 4, 5]] = r
                asdq = np.zeros((7, n + 4), order="F")
                asdq[:, :] = vec[:, None]
                bm, bp = O.rpt2(prm, ixy, 2, q1d, aux - 1.0, aux, aux - 1.0, asdq)
                got = (bm + bp)[:, 3]
                np.testing.assert_allclose(got[:6], lt * vec[:6], rtol=1e-12,
                                           atol=1e-12 * abs(lt) * np.abs(vec).max())
            asdq = np.zeros((7, n + 4), order="F")
            asdq[iu] = 1.0
            bm, bp = O.rpt2(prm, ixy, 1, q1d, aux - 1.0, aux, aux - 1.0, asdq)
            want = np.zeros(7)
            want[iu] = ut
            np.testing.assert_allclose((bm + bp)[:, 3], want, rtol=0, atol=1e-14)


def pin_rpt2_uses_the_receiving_cell(O):
    """The left-going fluctuation of edge i-1/2 enters cell i-1, the
    right-going one cell i, and each is split with the eigenstructure of the
    cell it enters (LeVeque 2002 sec. 21.3; rpt2_dclaw.f:16-17 imp = 1, 2):
    on a slice of varying states, edge i's output equals that of a uniform
    slice holding the receiving cell's state."""
    p, prm = _params(O)
    rng = np.random.default_rng(7)
    n = 8
    q1d = np.zeros((7, n + 4), order="F")
    h = rng.uniform(0.3, 2.0, n + 4)
    q1d[0], q1d[1], q1d[2] = h, h * rng.normal(0, 1, n + 4), h * rng.normal(0, 1, n + 4)
    q1d[3], q1d[4] = h * rng.uniform(0.4, 0.65, n + 4), 0.7 * p.rho_f * G * h
    aux = np.zeros((6, n + 4), order="F")
    asdq = np.asfortranarray(rng.normal(0, 1, (7, n + 4)))
    for imp in (1, 2):
        bm, bp = O.rpt2(prm, 1, imp, q1d, aux - 1.0, aux, aux - 1.0, asdq)
        for i in range(2, n + 2):
            c = i - 1 if imp == 1 else i
            uni = np.asfortranarray(np.repeat(q1d[:, c:c + 1], n + 4, axis=1))
            bm1, bp1 = O.rpt2(prm, 1, imp, uni, aux - 1.0, aux, aux - 1.0, asdq)
            assert np.array_equal(bm[:, i], bm1[:, i]) and np.array_equal(bp[:, i], bp1[:, i])


# ------------------------------------------------ the Riemann solver
def _solver(O, prm, **kw):
    base = dict(chiL=0.0, chiR=0.0, taudirL=1.0, taudirR=1.0, fsL=1.0, fsR=1.0, bL=0.0, bR=0.0,
                thetaL=0.0, thetaR=0.0)
    base.update(kw)
    for s in "LR":
        base["u" + s] = base["hu" + s] / base["h" + s]
        base["v" + s] = base["hv" + s] / base["h" + s]
        base["hm" + s] = base["m" + s] * base["h" + s]
    return O.riemann_solver(prm, 1, False, **base)


def pin_solver_waves_against_the_flow(O):
    """Flow to the right (left): the 1-wave (3-wave) travels against it and
    must carry exactly its own side's eigenvector ratios: hv/h = v, hm/h = m,
    pb/h = rho g varrho (theory.md:104-111, 147-150) and hchi/h = chi, while
    the contact's jumps of v, m, pb, chi join the other wave."""
    p, prm = _params(O)
    for sgn in (1.0, -1.0):
        hL, hR = 1.2, 0.9
        st = dict(hL=hL, hR=hR, huL=sgn * 0.5 * hL, huR=sgn * 0.55 * hR, hvL=0.3 * hL,
                  hvR=-0.2 * hR, mL=0.62, mR=0.5, pL=0.6 * 1000 * G * hL,
                  pR=0.8 * 1000 * G * hR, chiL=0.2, chiR=0.7)
        sw, fw = _solver(O, prm, **st)
        assert sw[0] < 0.0 < sw[2]
        k, s = (0, "L") if sgn > 0 else (2, "R")
        other = 2 - k
        m = st["m" + s]
        rho = m * p.rho_s + (1 - m) * p.rho_f
        varrho = (p.rho_f + 3 * rho) / (4 * rho)
        h = st["h" + s]
        ratios = [st["hv" + s] / h, m, rho * G * varrho, st["chi" + s]]
        assert abs(fw[0, k]) > 1e-3
        for row, want in zip((2, 3, 4, 5), ratios):
            assert fw[row, k] == pytest.approx(fw[0, k] * want, rel=1e-12, abs=1e-300), (sgn, row)
        # the contact's jump is not lost: it sits in the other wave
        m_o = st["m" + ("R" if s == "L" else "L")]
        assert abs(fw[3, other] - fw[0, other] * m_o) > 1e-6


def pin_solver_stationary_bed_step(O):
    """Constant discharge q over a bed step db, with the depth jump the
    solver assigns to the step (hR - hL = db g hbar / s1s2bar): the momentum
    flux difference equals the bed source (both use the same jump), so the
    mass and momentum f-waves vanish."""
    p, prm = _params(O)
    for hL, uL, db in [(1.0, 0.8, 0.05), (1.5, -0.6, -0.08), (0.7, 1.2, 0.02)]:
        q = hL * uL
        hR = hL
        for _ in range(200):
            uR = q / hR
            hbar = 0.5 * (hL + hR)
            s1s2bar = 0.25 * (uL + uR) ** 2 - G * hbar
            hR = hL + db * G * hbar / s1s2bar
        st = dict(hL=hL, hR=hR, huL=q, huR=q, hvL=0.0, hvR=0.0, mL=0.6, mR=0.6,
                  pL=0.7 * 1000 * G * hL, pR=0.7 * 1000 * G * hR, bL=0.0, bR=db)
        sw, fw = _solver(O, prm, **st)
        scale = G * hL * abs(db)
        assert np.abs(fw[0:2]).max() <= 1e-12 * scale, (hL, uL, db, fw[0:2])
        # and it is a real test: the same step under still water does make waves
        sw, fw = _solver(O, prm, **dict(st, huL=1e-3 * q, huR=1e-3 * q))
        assert np.abs(fw[0:2]).max() > 1e-4 * scale


# ------------------------------------------------ source rates (src2)
def _cell(h, hu, hv, m, pb):
    q = np.zeros((7, 5, 5), order="F")
    q[:, 2, 2] = [h, hu, hv, m * h, pb, 0.0, 0.0]
    return q, np.zeros((6, 5, 5), order="F")


def _rates(O, prm, q, aux, dt):
    """Richardson-extrapolated d/dt at t = 0 of h, hu, hv, hm, pb under src2
    (first-order one-sided differences at dt and dt/2)."""
    a = (O.src2(prm, q, aux, dt)[:5, 2, 2] - q[:5, 2, 2]) / dt
    b = (O.src2(prm, q, aux, 0.5 * dt)[:5, 2, 2] - q[:5, 2, 2]) / (0.5 * dt)
    return 2.0 * b - a


def pin_basal_friction_rates(O):
    """Moving material with effective stress and k = 0 (no pressure change
    from relaxation at t = 0): d(h v)/dt = -(tau_s + tau_f) v/(rho |v|) with
    tau_s = sigma_e tan(phi + psi), tau_f = 2 mu (1 - m) |v|/h
    (theory.md:127-135, 186-194), psi from the critical-state relation as in
    pin_dilatancy_rate; and the viscous decay alone (water, m = 0) is the
    exact exponential exp(-2 mu (1 - m) t/(rho h^2)) of that ODE, here also
    for m > 0 at lithostatic pressure (sigma_e = 0: no Coulomb part)."""
    for h, u, v, m, frac in [(1.0, 2.0, 1.0, 0.6, 0.4), (0.7, -3.0, 0.5, 0.62, 0.7),
                             (1.5, 0.4, -2.5, 0.5, 0.0)]:
        p, prm = _params(O, kref=0.0)
        rho = m * p.rho_s + (1 - m) * p.rho_f
        pb = frac * rho * G * h
        q, aux = _cell(h, h * u, h * v, m, pb)
        sig = rho * G * h - pb
        speed = math.hypot(u, v)
        gdot = 2 * speed / h
        N = p.mu * gdot / (p.rho_s * gdot ** 2 * p.delta ** 2 + sig)
        tanpsi = (m - p.m_crit / (1 + math.sqrt(N))) / p.c1
        tau_s = sig * math.tan(math.radians(p.phi_deg) + math.atan(tanpsi))
        tau_f = 2 * p.mu * (1 - m) * speed / h
        r = _rates(O, prm, q, aux, 1e-4)
        k = -(tau_s + tau_f) / rho / speed
        assert r[1] == pytest.approx(k * u, rel=1e-7), (h, u, v, m)
        assert r[2] == pytest.approx(k * v, rel=1e-7), (h, u, v, m)
    for m in (0.0, 0.45):
        p, prm = _params(O, kref=0.0)
        rho = m * p.rho_s + (1 - m) * p.rho_f
        h, u, dt = 0.3, 1.5, 2.0
        q, aux = _cell(h, h * u, 0.0, m, rho * G * h)
        out = O.src2(prm, q, aux, dt)
        want = h * u * math.exp(-2 * p.mu * (1 - m) * dt / (rho * h * h))
        assert out[1, 2, 2] == pytest.approx(want, rel=1e-13), m


def pin_pressure_relaxation_rate(O):
    """Still material: dpb/dt = zeta D with zeta = 3/(2 alpha h) +
    g rho_f (rho - rho_f)/(4 rho), D = -(2k/(h mu))(pb - rho_f g h),
    1/alpha = m (rho g h - pb + sigma_0)/a (theory.md:142-174); m = m_ref so
    k = k_ref.  Also dh/dt = D (rho - rho_f)/rho and d(hm)/dt = -D m rho_f/rho
    (theory.md:131-139)."""
    for kref, frac, h, dm in [(1e-8, 0.3, 1.0, 0.0), (3e-9, 1.6, 0.6, 0.0),
                              (1e-8, 0.9, 2.0, 0.0), (1e-8, 0.5, 1.0, 0.04),
                              (1e-9, 1.3, 0.8, -0.06)]:
        p, prm = _params(O, kref=kref)
        m = p.mref + dm
        rho = m * p.rho_s + (1 - m) * p.rho_f
        pb = frac * p.rho_f * G * h
        q, aux = _cell(h, 0.0, 0.0, m, pb)
        ainv = m * (rho * G * h - pb + p.sigma_0) / p.alpha_c
        zeta = 1.5 * ainv / h + G * p.rho_f * (rho - p.rho_f) / (4 * rho)
        # permeability falls e-fold per 0.04 of solid fraction above m_ref
        # (Iverson and George 2014, eq. 2.9; digclaw_module.f90:431)
        k = kref * math.exp(-dm / 0.04)
        D = -(2 * k / (h * p.mu)) * (pb - p.rho_f * G * h)
        r = _rates(O, prm, q, aux, 1e-5)
        assert r[4] == pytest.approx(zeta * D, rel=1e-6), (kref, frac, h, dm)
        assert r[0] == pytest.approx(D * (rho - p.rho_f) / rho, rel=1e-6)
        assert r[3] == pytest.approx(-D * m * p.rho_f / rho, rel=1e-6)
        assert r[1] == 0.0 and r[2] == 0.0


def pin_dilatancy_rate(O):
    """Moving material at lithostatic pressure (no effective stress, so no
    Coulomb friction) with k = 0 (no relaxation): dpb/dt =
    -3 |v| tan(psi)/(alpha h) with c1 tan(psi) = m - m_crit/(1 + sqrt(N)),
    N = mu gdot/(rho_s gdot^2 delta^2 + sigma_e), gdot = 2|v|/h
    (theory.md:142, 196-208; the shear rate of setvars,
    digclaw_module.f90:394-491), 1/alpha = m sigma_0/a."""
    for h, u, v, m in [(1.0, 1.0, 0.0, 0.62), (0.5, -0.6, 0.8, 0.60), (2.0, 0.2, 0.3, 0.45)]:
        p, prm = _params(O, kref=0.0)
        rho = m * p.rho_s + (1 - m) * p.rho_f
        q, aux = _cell(h, h * u, h * v, m, rho * G * h)
        speed = math.hypot(u, v)
        gdot = 2 * speed / h
        N = p.mu * gdot / (p.rho_s * gdot ** 2 * p.delta ** 2 + 0.0)
        tanpsi = (m - p.m_crit / (1 + math.sqrt(N))) / p.c1
        ainv = m * p.sigma_0 / p.alpha_c
        want = -3.0 * speed * tanpsi * ainv / h
        r = _rates(O, prm, q, aux, 1e-6)
        if want > 0.0:
            continue        # compression above lithostatic is clipped by qfix
        assert r[4] == pytest.approx(want, rel=1e-6), (h, u, v, m)


def pin_rpn2_fluctuations(O):
    """rpn2 on a slice with one jump (flat bed, both sides wet, moving):
    amdq + apdq is the flux difference of every conserved component -- mass
    h u_n, normal momentum h u_n^2 + g h^2/2, transverse momentum h u_n u_t,
    hm u_n, h chi u_n (the f-wave splitting, LeVeque 2002 sec. 15.5, through
    rpn2's map of 3 solver waves to 5, rpn2_dclaw.f:276-379); amdq holds the
    left-going part.  Both sweep directions."""
    p, prm = _params(O)
    for ixy in (1, 2):
        for sgn in (1.0, -1.0):
            n = 6
            q1d = np.zeros((7, n + 4), order="F")
            iu, iv = (1, 2) if ixy == 1 else (2, 1)
            hL, hR = 1.1, 0.8
            left = dict(h=hL, un=sgn * 0.6, ut=0.4, m=0.6, chi=0.2, pf=0.6)
            right = dict(h=hR, un=sgn * 0.7, ut=-0.3, m=0.52, chi=0.6, pf=0.8)
            for i in range(n + 4):
                st = left if i < 5 else right
                h = st["h"]
                q1d[0, i], q1d[iu, i], q1d[iv, i] = h, h * st["un"], h * st["ut"]
                q1d[3, i], q1d[4, i], q1d[5, i] = h * st["m"], st["pf"] * 1000 * G * h, h * st["chi"]
            aux = np.zeros((6, n + 4), order="F")
            aux[3] = 1.0             # fsphi
            aux[4] = aux[5] = 1.0    # taudir
            _, fwave, s, amdq, apdq = O.rpn2(prm, ixy, q1d, aux)
            e = 5                    # the edge between cells 4 and 5 (0-based)
            tot = amdq[:, e] + apdq[:, e]

            def flux(st):
                h, un, ut = st["h"], st["un"], st["ut"]
                return np.array([h * un, h * un * un + 0.5 * G * h * h, h * un * ut,
                                 h * st["m"] * un, h * st["chi"] * un])
            want = flux(right) - flux(left)
            got = tot[[0, iu, iv, 3, 5]]
            np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13)
            left_going = sum(fwave[:, k, e] for k in range(5) if s[k, e] < 0)
            np.testing.assert_allclose(amdq[:, e], left_going, rtol=1e-15, atol=0)
            assert s[0, e] < 0 < s[4, e]
            # other edges: no jump, no fluctuation
            assert not amdq[:, 3].any() and not apdq[:, 3].any()


def pin_static_heap_threshold(O):
    """A uniform layer at rest on a plane of slope angle theta (coordinate
    system 1: bed source g h tan(theta)) stays exactly at rest while the
    driving stress is below the Coulomb resistance and starts to slide above
    it: tan(theta_c) = sigma_e tan(phi + psi)/(rho g h), sigma_e = rho g h -
    pb, c1 tan(psi) = m - m_crit at zero shear (theory.md:186-208).  This runs
    calc_taudir, the solver's static branch and fsphi together
    (riemannsolvers_dclaw.f:308-341, digclaw_module.f90:500-719)."""
    p = config0_params()
    m, frac = 0.62, 0.5
    rho = m * p.rho_s + (1 - m) * p.rho_f
    sig = rho * G - frac * p.rho_f * G
    psi = math.atan((m - p.m_crit) / p.c1)
    theta_c = math.degrees(math.atan(sig * math.tan(math.radians(p.phi_deg) + psi) / (rho * G)))
    prm = O.params_from(p)
    mx, my, dx = 30, 12, 1.0
    x = (np.arange(mx) + 0.5) * dx
    for dtheta, moves in [(-1.5, False), (1.5, True), (-8.0, False)]:
        th = math.radians(theta_c + dtheta)
        b = -math.tan(th) * x[:, None] * np.ones((1, my))
        pt = W.make_patch(b, np.ones((mx, my)), m, frac, p, dx, dx)
        q, _, _, _, _ = O.run(prm, pt.q, pt.aux, dx, dx, 40, 1e-3, 0.9, 0.01)
        vmax = max(np.abs(q[1, 2:-2, 2:-2]).max(), np.abs(q[2, 2:-2, 2:-2]).max())
        if moves:
            assert vmax > 1e-3, (theta_c, dtheta, vmax)
        else:
            assert vmax == 0.0, (theta_c, dtheta, vmax)


def pin_calc_taudir_still_cell(O):
    """A still layer of uniform depth under a planar free surface
    eta = a x + c y (anisotropic cells): the driving force is -g h grad(eta), so the friction direction
    stored by calc_taudir is +grad(eta)/|grad(eta)| scaled by (dx, dy)
    (digclaw_module.f90:641-713), with fsphi = 1."""
    p, prm = _params(O)
    dx, dy = 2.0, 3.0
    mx, my = 6, 5
    for a, c in [(0.05, -0.02), (-0.01, 0.03), (0.0, 0.04)]:
        x = (np.arange(-1, mx + 3) - 0.5) * dx
        y = (np.arange(-1, my + 3) - 0.5) * dy
        X, Y = np.meshgrid(x, y, indexing="ij")
        eta = 10.0 + a * X + c * Y
        b = eta - 2.0              # uniform depth: the averaged depths agree
        q = np.zeros((7, mx + 4, my + 4), order="F")
        q[0] = eta - b
        q[3] = 0.6 * q[0]
        q[4] = 0.5 * 1000 * G * q[0]
        aux = np.zeros((6, mx + 4, my + 4), order="F")
        aux[0] = b
        out = O.calc_taudir(prm, q, aux, dx, dy)
        g = math.hypot(a, c)
        np.testing.assert_allclose(out[4, 3, 3], dx * a / g, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(out[5, 3, 3], dy * c / g, rtol=1e-12, atol=1e-14)
        assert out[3, 3, 3] == 1.0


def pin_fe4quad_rates(O):
    """src2method 2 (mp_update_FE_4quad) integrates the same source terms on
    the rho h = const manifold: at vanishing dt its rates equal theory.md's
    (dpb/dt = zeta D - 3 |v| tan(psi)/(alpha h), dh/dt = D (rho - rho_f)/rho,
    d(hm)/dt = -D m rho_f/rho, with sigma_0 of alphamethod 1,
    theory.md:139-190) in each of the four quadrants of (m, pb) around the
    critical point; and a still cell (vnorm = 0, where the reference reads an
    unset time step, mp_update.f90:226-254) comes out as NaN."""
    seen = set()
    for m, frac, speed in [(0.50, 1.2, 1.0), (0.66, 1.2, 1.0), (0.66, 0.6, 1.0), (0.50, 0.6, 1.0),
                           (0.55, 1.1, 3.0), (0.68, 0.3, 0.5)]:
        p, prm = _params(O, kref=1e-9, src2method=2, alphamethod=1)
        h = 1.0
        rho = m * p.rho_s + (1 - m) * p.rho_f
        pb = frac * p.rho_f * G * h
        u, v = 0.8 * speed, 0.6 * speed
        q, aux = _cell(h, h * u, h * v, m, pb)
        sig = rho * G * h - pb
        sig0 = 0.5 * p.alpha_c * p.rho_f * (p.rho_s - p.rho_f) * G * h / rho
        ainv = m * (sig + sig0) / p.alpha_c
        gdot = 2 * speed / h
        N = p.mu * gdot / (p.rho_s * gdot ** 2 * p.delta ** 2 + sig)
        tanpsi = (m - p.m_crit / (1 + math.sqrt(N))) / p.c1
        k = p.kref * math.exp(-(m - p.mref) / 0.04)
        D = -(2 * k / (h * p.mu)) * (pb - p.rho_f * G * h)
        zeta = 1.5 * ainv / h + G * p.rho_f * (rho - p.rho_f) / (4 * rho)
        r = _rates(O, prm, q, aux, 1e-6)
        assert r[4] == pytest.approx(zeta * D - 3 * speed * tanpsi * ainv / h, rel=1e-5), (m, frac)
        assert r[0] == pytest.approx(D * (rho - p.rho_f) / rho, rel=1e-5), (m, frac)
        assert r[3] == pytest.approx(-D * m * p.rho_f / rho, rel=1e-5), (m, frac)
        seen.add((m > 0.6, frac > 1.0))
    assert len(seen) == 4
    p, prm = _params(O, src2method=2, alphamethod=1)
    q, aux = _cell(1.0, 0.0, 0.0, 0.6, 0.5 * 1000 * G)
    assert math.isnan(O.src2(prm, q, aux, 0.01)[4, 2, 2])


def pin_curvature_running_gravity(O):
    """src2 with curvature on a parabolic bed b = kappa x^2 / 2 (b_xx = kappa,
    the centred second difference of a quadratic is exact) under a uniform
    flow u: each wet cell's curvature acceleration is u^2 kappa, and the
    gravity src2 uses is set once and raised by every wet cell in loop order
    (src2.f90:87, 152-156: j outer, i inner).  A pore pressure above
    lithostatic is clipped by qfix to rho gz h, which makes cell n of the
    sweep show pb = rho h (g + n u^2 kappa) (src2method -1: nothing else
    changes pb)."""
    p, prm = _params(O, src2method=-1, curvature=1)
    mx, my, dx, dy = 10, 6, 2.0, 1.5
    kappa, u, h, m = 0.05, 2.0, 1.0, 0.6
    x = (np.arange(-2, mx + 2) + 0.5) * dx
    aux = np.zeros((3, mx + 4, my + 4), order="F")
    aux[0] = (0.5 * kappa * x * x)[:, None]
    q = np.zeros((7, mx + 4, my + 4), order="F")
    q[0], q[1], q[3], q[4] = h, h * u, m * h, 1e7
    out = O.src2(prm, q, aux, 1e-3, dx=dx, dy=dy)
    rho = m * p.rho_s + (1 - m) * p.rho_f
    gacc = u * u * kappa
    for j in range(1, my + 1):
        for i in range(1, mx + 1):
            n = (j - 1) * mx + i
            assert out[4, i + 1, j + 1] == pytest.approx(rho * (G + n * gacc) * h, rel=1e-12), (i, j)


def pin_speed_limit(O):
    """b4step2: a cell faster than speed_limit is slowed to exactly the limit
    with its direction kept; slower cells are untouched."""
    p, prm = _params(O)
    q = np.zeros((7, 6, 6), order="F")
    q[0] = 1.0
    q[3] = 0.6
    q[4] = 0.5 * 1000 * G
    q[1, 2, 2], q[2, 2, 2] = 30.0, -40.0          # |v| = 50 > 20
    q[1, 3, 3], q[2, 3, 3] = 12.0, 9.0            # |v| = 15
    aux = np.zeros((6, 6, 6), order="F")
    out, _ = O.b4step2(prm, q, aux, 1.0, 1.0)
    hu, hv = out[1, 2, 2], out[2, 2, 2]
    assert math.hypot(hu, hv) == pytest.approx(p.speed_limit, rel=1e-15)
    assert hu / hv == pytest.approx(30.0 / -40.0, rel=1e-15)
    assert (out[1, 3, 3], out[2, 3, 3]) == (12.0, 9.0)


def pin_still_impermeable_cell_keeps_pressure(O):
    """k = 0 and no motion: nothing drives pb (D = 0, no dilatancy), so any
    admissible pb in [0, rho g h] is kept, also far below hydrostatic; and
    for random states src2 leaves pb within [0, rho g h] (qfix_cmass
    bounds, digclaw_module.f90:353-387)."""
    p, prm = _params(O, kref=0.0)
    h, m = 1.0, 0.6
    rho = m * p.rho_s + (1 - m) * p.rho_f
    for frac in (0.0, 0.05, 0.3, 0.9, 1.0):
        pb = frac * rho * G * h
        q, aux = _cell(h, 0.0, 0.0, m, pb)
        out = O.src2(prm, q, aux, 0.5)
        assert out[4, 2, 2] == pytest.approx(pb, rel=1e-14, abs=1e-9), frac
    p, prm = _params(O)
    rng = np.random.default_rng(2)
    for _ in range(200):
        h = rng.uniform(0.01, 3.0)
        m = rng.uniform(0.3, 0.7)
        rho = m * p.rho_s + (1 - m) * p.rho_f
        q, aux = _cell(h, h * rng.normal(0, 2), h * rng.normal(0, 2), m,
                       rng.uniform(0, 1) * rho * G * h)
        out = O.src2(prm, q, aux, rng.uniform(1e-3, 0.1))
        h2, m2 = out[0, 2, 2], out[3, 2, 2] / out[0, 2, 2]
        rho2 = m2 * p.rho_s + (1 - m2) * p.rho_f
        assert 0.0 <= out[4, 2, 2] <= rho2 * G * h2 * (1 + 1e-14)


# ------------------------------------------------ Thacker's bowl (tests/bowl-slosh)
def _bowl_run(O, n, T, dry_tolerance, cfl=0.75, dt0=1e-4):
    """workloads.bowl_slosh on n x n cells to time T with the reference
    case's time-step rule (cfl_desired 0.75, dt_initial 1e-4: setrun.py:155-167)."""
    pt = W.bowl_slosh(n, dry_tolerance)
    prm = O.params_from(pt.params)
    q, aux, t, dt = pt.q, pt.aux, 0.0, dt0
    while t < T * (1 - 1e-14):
        dt = min(dt, T - t)
        q = O.bc_fill(q, (1, 1, 1, 1))
        q, aux, c = O.stepgrid(prm, q, aux, pt.dx, pt.dy, dt)
        t += dt
        dt = dt * cfl / c if c > 0 else 2 * dt
    return pt, q[:, 2:-2, 2:-2]


def pin_bowl_slosh_thacker(O):
    """Clear water (m = 0, hydrostatic pb) sloshing in a paraboloid bowl
    converges to Thacker's planar-surface solution: over a quarter period
    the mean depth error and the velocity error of the deep cells fall with
    the grid (first order is what a moving shoreline allows), mass is kept
    when the dry tolerance is negligible, and the source terms leave m = 0
    and pb = rho_f g h exactly (zero effective stress: no Coulomb friction,
    no relaxation, no dilatancy).  At the reference case's own 41 x 41 grid
    and dry tolerance 1e-3, the state at its output time t = 1 stays close
    to the solution and qfix's zeroing of cells thinner than the tolerance
    (digclaw_module.f90 qfix) only removes mass."""
    from helpers import bowl_errors
    T = 0.25 * 2 * math.pi / math.sqrt(2 * G * W.BOWL_H0)
    eh, ev = [], []
    for n in (40, 80, 160):
        pt, q = _bowl_run(O, n, T, 1e-6)
        q0 = pt.q[:, 2:-2, 2:-2]
        assert not q[3].any()
        assert np.abs(q[4] - pt.params.rho_f * G * q[0]).max() <= 1e-10 * 1000.0 * G * 0.1
        assert abs(q[0].sum() / q0[0].sum() - 1) < 1e-5
        a, b = bowl_errors(q, n, T)
        eh.append(a)
        ev.append(b)
    assert eh[0] > eh[1] > eh[2] and eh[0] / eh[2] > 3 and eh[2] < 2e-4, eh
    assert ev[0] > ev[1] > ev[2] and ev[0] / ev[2] > 2.5 and ev[2] < 0.06, ev
    pt, q = _bowl_run(O, 41, 1.0, 1e-3)
    a, b = bowl_errors(q, 41, 1.0)
    assert a < 1e-3 and b < 0.25, (a, b)
    loss = q[0].sum() / pt.q[0, 2:-2, 2:-2].sum() - 1
    assert -0.03 < loss <= 1e-12, loss


PINS = [pin_rpt2_gravity_eigenvectors, pin_rpt2_uses_the_receiving_cell,
        pin_solver_waves_against_the_flow,
        pin_solver_stationary_bed_step, pin_rpn2_fluctuations, pin_basal_friction_rates,
        pin_pressure_relaxation_rate, pin_dilatancy_rate, pin_speed_limit,
        pin_still_impermeable_cell_keeps_pressure, pin_static_heap_threshold,
        pin_calc_taudir_still_cell, pin_fe4quad_rates, pin_curvature_running_gravity,
        pin_bowl_slosh_thacker, pin_second_order_1d, pin_second_order_2d]


@pytest.mark.parametrize("pin", PINS, ids=[f.__name__[4:] for f in PINS])
def test_pin(oracle_mod, pin):
    pin(oracle_mod)
