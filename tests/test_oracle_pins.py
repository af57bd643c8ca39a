"""Pins of the CPU oracle against known answers, closed forms and invariants.

Each check is chosen so that a plausible transcription mistake (a dropped
term, a sign, a swapped index or operand) fails it; none of them retypes the
oracle's own formula.  Sources: a known answer of riemanntype from a run of
the reference's compiled Fortran, the reference's eigenstructure tests
(tests/eigenstructure_check/check.py:41-121), textbook shallow-water
identities (f-wave sums, the transverse Jacobian, lake at rest), and the
closed-form solutions of the source sub-steps.
"""
import math

import numpy as np
import pytest

from dclaw_amd import workloads as W
from dclaw_amd.params import DigParams, config0_params
from helpers import slope_patch


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def P(O, **kw):
    p = config0_params()
    for k, v in kw.items():
        setattr(p, k, v)
    return O.params_from(p)


# ---------------------------------------------------------------- riemanntype
def test_riemanntype_known_answer(O):
    # riemanntype(hL=2, hR=1, uL=0, uR=0) with g = 9.81, drytol = 1e-3, run
    # through the reference's own Fortran (compiled reference probe).
    hm, s1m, s2m, r1, r2 = O.riemanntype(2.0, 1.0, 0.0, 0.0, 1, 1e-3, 9.81)
    assert hm == 1.4889054787888178
    assert s1m == -2.6065091102140361
    assert s2m == 5.0370928540220152
    assert (r1, r2) == (True, False)


@pytest.mark.parametrize("hL,hR,uL,uR", [(1.0, 1.5, -2.0, 3.0), (0.3, 0.2, -1.0, 1.5)])
def test_riemanntype_two_rarefaction_closed_form(O, hL, hR, uL, uR):
    g = 9.81
    hm, s1m, s2m, r1, r2 = O.riemanntype(hL, hR, uL, uR, 1, 1e-3, g)
    assert r1 and r2
    # exact middle depth of the two-rarefaction solution (LeVeque 2002, 13.56)
    expect = (uL - uR + 2 * (math.sqrt(g * hL) + math.sqrt(g * hR))) ** 2 / (16 * g)
    assert hm == pytest.approx(expect, rel=1e-14)
    assert s1m == pytest.approx(uL + 2 * math.sqrt(g * hL) - 3 * math.sqrt(g * hm), rel=1e-13)


def test_riemanntype_dry_side(O):
    g = 9.81
    hm, s1m, s2m, r1, r2 = O.riemanntype(2.0, 0.0, 1.0, 0.0, 1, 1e-3, g)
    assert hm == 0.0
    # dry-bed front speed u_L + 2 sqrt(g h_L)
    assert s1m == pytest.approx(1.0 + 2 * math.sqrt(2 * g), rel=1e-15)
    assert (r1, r2) == (True, False)


def test_riemanntype_two_shock_one_newton_step(O):
    # colliding streams: one Newton step on sqrt(h) from h_max; the result
    # must exceed both depths and be a better root than the start point.
    g = 9.81
    hL, hR, uL, uR = 1.0, 1.2, 3.0, -3.0
    hm, s1m, s2m, r1, r2 = O.riemanntype(hL, hR, uL, uR, 1, 1e-3, g)
    assert not r1 and not r2
    F = lambda h: (uR - uL) + (h - hL) * math.sqrt(0.5 * g * (1 / h + 1 / hL)) \
        + (h - hR) * math.sqrt(0.5 * g * (1 / h + 1 / hR))
    assert hm > max(hL, hR)
    assert abs(F(hm)) < abs(F(max(hL, hR)))


# ---------------------------------------------------------------- solver
def _solver_case(O, prm, **kw):
    base = dict(hL=1.2, hR=0.9, huL=0.6, huR=0.45, hvL=0.12, hvR=-0.09, bL=0.0, bR=0.0,
                mL=0.6, mR=0.58, chiL=0.0, chiR=0.0, taudirL=1.0, taudirR=1.0, fsL=1.0, fsR=1.0)
    base.update(kw)
    g = 9.81
    rho = lambda m: m * 2700 + (1 - m) * 1000
    base.setdefault("uL", base["huL"] / base["hL"])
    base.setdefault("uR", base["huR"] / base["hR"])
    base.setdefault("vL", base["hvL"] / base["hL"])
    base.setdefault("vR", base["hvR"] / base["hR"])
    base.setdefault("hmL", base["mL"] * base["hL"])
    base.setdefault("hmR", base["mR"] * base["hR"])
    base.setdefault("pL", 0.7 * 1000 * g * base["hL"])
    base.setdefault("pR", 0.75 * 1000 * g * base["hR"])
    sw, fw = O.riemann_solver(prm, 1, False, **base)
    return base, sw, fw


def test_solver_fwave_sums_flat_bed(O):
    """On a flat bed with moving material, the f-waves split the flux
    differences exactly: mass flux, momentum flux hu^2 + g h^2/2, and the
    advective fluxes of hv and hm (riemannsolvers_dclaw.f:420-450)."""
    prm = P(O)
    x, sw, fw = _solver_case(O, prm)
    g = 9.81
    hL, hR, uL, uR = x["hL"], x["hR"], x["uL"], x["uR"]
    assert fw[0].sum() == pytest.approx(x["huR"] - x["huL"], rel=1e-13, abs=1e-15)
    mom = lambda h, u: h * u * u + 0.5 * g * h * h
    assert fw[1].sum() == pytest.approx(mom(hR, uR) - mom(hL, uL), rel=1e-12)
    assert fw[2].sum() == pytest.approx(x["hvR"] * uR - x["hvL"] * uL, rel=1e-12)
    assert fw[3].sum() == pytest.approx(x["hmR"] * uR - x["hmL"] * uL, rel=1e-12)
    # middle wave carries only normal momentum after the fold (:513-516)
    assert fw[0, 1] == 0.0 and all(fw[k, 1] == 0.0 for k in (2, 3, 4, 5))
    assert sw[0] < sw[1] < sw[2]


def test_solver_lake_at_rest(O):
    """h + b constant, u = v = 0: the bed-step source balances the hydrostatic
    jump and every f-wave vanishes (well-balanced)."""
    prm = P(O)
    for m in (0.0, 0.6):
        x, sw, fw = _solver_case(O, prm, hL=2.0, hR=1.5, huL=0.0, huR=0.0, hvL=0.0, hvR=0.0,
                                 bL=0.0, bR=0.5, mL=m, mR=m,
                                 pL=1000 * 9.81 * 2.0, pR=1000 * 9.81 * 1.5)
        assert np.abs(fw).max() <= 1e-12 * 9.81 * 2.0 ** 2


def test_solver_static_branch(O):
    """Still material on a slope steeper than nothing but weaker than the
    friction: the static branch zeros mass, pressure and hm waves."""
    prm = P(O)
    x, sw, fw = _solver_case(O, prm, hL=1.0, hR=1.0, huL=0.0, huR=0.0, hvL=0.0, hvR=0.0,
                             bL=0.02, bR=0.0, pL=0.5 * 1000 * 9.81, pR=0.5 * 1000 * 9.81)
    assert fw[0].sum() == 0.0 and fw[3].sum() == 0.0 and fw[4].sum() == 0.0
    assert np.abs(fw[1]).sum() == 0.0


# ---------------------------------------------------------------- rpt2
def test_rpt2_reproduces_transverse_jacobian(O):
    """With no gating, bmasdq + bpasdq = B(q) asdq on (h, hu, hv) where B is
    the textbook Jacobian of the y-flux (hv, huv, hv^2 + g h^2/2)."""
    prm = P(O)
    n = 5 + 4
    q1d = np.zeros((7, n), order="F")
    h, u, v = 1.3, 0.7, -0.4
    q1d[0], q1d[1], q1d[2], q1d[3], q1d[4] = h, h * u, h * v, 0.6 * h, 0.7 * 1000 * 9.81 * h
    aux = np.zeros((6, n), order="F")
    asdq = np.zeros((7, n), order="F")
    rng = np.random.default_rng(1)
    asdq[:6] = rng.standard_normal((6, n))
    bm, bp = O.rpt2(prm, 1, 2, q1d, aux - 1.0, aux, aux - 1.0, asdq)
    g = 9.81
    B = np.array([[0, 0, 1], [-u * v, v, u], [-v * v + g * h, 0, 2 * v]])
    for i in range(1, n):
        want = B @ asdq[:3, i]
        got = bm[:3, i] + bp[:3, i]
        assert np.allclose(got, want, rtol=1e-13, atol=1e-13)


def test_rpt2_dry_and_blocked(O):
    prm = P(O)
    n = 6
    q1d = np.zeros((7, n), order="F")
    aux = np.zeros((6, n), order="F")
    asdq = np.ones((7, n), order="F")
    bm, bp = O.rpt2(prm, 1, 1, q1d, aux, aux, aux, asdq)
    assert not bm.any() and not bp.any()
    q1d[0] = 1.0
    bm, bp = O.rpt2(prm, 1, 1, q1d, aux + 5.0, aux, aux + 5.0, asdq)  # walls above and below
    assert not bm.any() and not bp.any()


# ---------------------------------------------------------------- limiter
@pytest.mark.parametrize("ratio,phi", [(-1.0, 0.0), (0.25, 0.5), (1.0, 1.0), (2.0, 1.5),
                                       (3.0, 2.0), (10.0, 2.0)])
def test_mc_limiter_closed_form(O, ratio, phi):
    """Collinear upwind waves with ratio r: MC gives max(0, min((1+r)/2, 2, 2r))."""
    prm = P(O)
    n = 3 + 4
    fw = np.zeros((7, 5, n), order="F")
    s = np.full((5, n), 1.0, order="F")
    w = np.array([1.0, 2.0, -0.5, 0.3, 4.0, 0.0, 0.0])
    for i in range(n):
        fw[:, 0, i] = w * (ratio ** (-i))    # wave(i-1) = ratio * wave(i)
    out = O.limiter(prm, fw, s)
    i = 3  # an edge with an upwind neighbour
    assert out[:, 0, i] == pytest.approx(phi * fw[:, 0, i], rel=1e-15, abs=0)


# ---------------------------------------------------------------- qfix / setvars
def test_qfix_admissibility(O):
    prm = P(O)
    q, prim = O.qfix(prm, [2.0, 1.0, -1.0, 2.5, 1e9, -0.3])
    h, hu, hv, hm, p, hchi = q
    assert hm == h * 1.0 and hchi == 0.0           # m, chi clamped to [0, 1]
    assert p == 2700 * 9.81 * 2.0                  # p clamped to rho g h with m = 1
    q, prim = O.qfix(prm, [5e-4, 1.0, 1.0, 1e-4, 3.0, 0.0])
    assert not np.any(q) and not np.any(prim)       # dry: everything zero


def test_setvars_special_values(O):
    prm = P(O)
    g = 9.81
    # m = mref: k = kref exactly; u = v = 0: m_eq = m_crit, shear 0
    r = O.setvars(prm, 1.0, 0.0, 0.0, 0.6, 1000 * g, 0.0, g)
    assert r["kperm"] == 1e-10
    assert r["m_eq"] == 0.64
    rho = 0.6 * 2700 + 0.4 * 1000
    sig = rho * g - 1000 * g
    assert r["alphainv"] == pytest.approx(0.6 * (sig + 1000.0) / 0.01, rel=1e-15)
    # tau = sig tan(phi + atan(c1 (m - m_eq))) * (1 + tanh(100 (m - 0.1)))/2
    want = sig * math.tan(math.radians(40) + math.atan(0.6 - 0.64)) * 0.5 * (1 + math.tanh(50.0))
    assert r["tau"] == pytest.approx(want, rel=1e-14)
    # lithostatic pressure: no effective stress, no friction
    r = O.setvars(prm, 1.0, 0.3, 0.0, 0.6, rho * g, 0.0, g)
    assert r["tau"] == 0.0


# ---------------------------------------------------------------- src2
def _one_cell(O, p, h, hu, hv, hm, pb):
    q = np.zeros((7, 5, 5), order="F")
    q[:, 2, 2] = [h, hu, hv, hm, pb, 0.0, 0.0]
    aux = np.zeros((6, 5, 5), order="F")
    return q, aux


def test_src2_water_viscous_decay(O):
    """Pure water (m = 0, hydrostatic pb): no Coulomb friction, no dilatancy,
    no relaxation; momentum decays as exp(-2 mu dt / (rho_f h^2))."""
    prm = P(O)
    h, u, dt = 0.5, 1.3, 0.01
    q, aux = _one_cell(O, prm, h, h * u, 0.0, 0.0, 1000 * 9.81 * h)
    out = O.src2(prm, q, aux, dt, mbc=2)
    want = h * u * math.exp(-2 * 0.005 * dt / (h * 1000 * h))
    assert out[1, 2, 2] == pytest.approx(want, rel=1e-14)
    assert out[0, 2, 2] == h and out[2, 2, 2] == 0.0


def test_src2_coulomb_stop_and_direction(O):
    prm = P(O)
    h = 1.0
    # slow, heavily loaded material stops exactly within one step
    q, aux = _one_cell(O, prm, h, 0.01, 0.02, 0.6 * h, 0.2 * 1000 * 9.81 * h)
    out = O.src2(prm, q, aux, 0.1, mbc=2)
    assert out[1, 2, 2] == 0.0 and out[2, 2, 2] == 0.0
    # fast material keeps its direction
    q, aux = _one_cell(O, prm, h, 3.0, -4.0, 0.6 * h, 0.2 * 1000 * 9.81 * h)
    out = O.src2(prm, q, aux, 1e-3, mbc=2)
    assert out[1, 2, 2] / out[2, 2, 2] == pytest.approx(3.0 / -4.0, rel=1e-14)
    assert 0 < math.hypot(out[1, 2, 2], out[2, 2, 2]) < 5.0


def test_src2_dry_and_negative_hm(O):
    prm = P(O)
    q, aux = _one_cell(O, prm, 5e-4, 1.0, 1.0, 1e-4, 3.0)
    out = O.src2(prm, q, aux, 0.01, mbc=2)
    assert not out[:6, 2, 2].any()
    q, aux = _one_cell(O, prm, 1.0, 1.0, 1.0, -0.1, 3.0)
    out = O.src2(prm, q, aux, 0.01, mbc=2)
    assert np.array_equal(out, q)                  # hm < 0: cell left untouched


def test_src2_pressure_relaxes_toward_hydrostatic(O):
    prm = P(O)
    h = 1.0
    g = 9.81
    for frac in (0.2, 0.95):
        q, aux = _one_cell(O, prm, h, 0.0, 0.0, 0.6 * h, frac * 2054.0 * g * h)
        out = O.src2(prm, q, aux, 5.0, mbc=2)
        p_eq = 1000 * g * out[0, 2, 2]
        assert abs(out[4, 2, 2] - p_eq) < abs(q[4, 2, 2] - 1000 * g * h)
        rho = out[3, 2, 2] / out[0, 2, 2] * 2700 + (1 - out[3, 2, 2] / out[0, 2, 2]) * 1000
        assert 0.0 <= out[4, 2, 2] <= rho * g * out[0, 2, 2]


# ---------------------------------------------------------------- calc_taudir
def test_calc_taudir_moving_and_still(O):
    prm = P(O)
    pt = slope_patch(8, 6, slope=0.0, h0=0.0)
    q = pt.q.copy(order="F")
    q[0, 2:-2, 2:-2] = 1.0
    q[3, 2:-2, 2:-2] = 0.6
    q[1, 4, 4], q[2, 4, 4] = 3.0, 4.0
    flat = pt.aux.copy(order="F")
    flat[0] = 0.0
    aux = O.calc_taudir(prm, q, flat, 2.0, 3.0)
    # moving cell: direction cosines scaled by dx, dy, opposing the motion
    assert aux[4, 4, 4] == pytest.approx(-2.0 * 3.0 / 5.0, rel=1e-15)
    assert aux[5, 4, 4] == pytest.approx(-3.0 * 4.0 / 5.0, rel=1e-15)
    # still cell on a flat free surface: no net force -> (dx, dy), fsphi = 1
    assert (aux[4, 6, 3], aux[5, 6, 3], aux[3, 6, 3]) == (2.0, 3.0, 1.0)
    # the outermost ghost ring is never touched
    assert not aux[3:6, 0, :].any() and not aux[3:6, :, -1].any()


# ---------------------------------------------------------------- patch level
def test_step2_conserves_mass_with_walls(O):
    """Without src2, the conservative update with wall boundaries keeps the
    total depth to rounding (f-waves telescope; walls carry no flux)."""
    pt = slope_patch(30, 20, slope=0.5, phi_deg=20.0)
    prm = O.params_from(pt.params)
    q = O.bc_fill(pt.q, (3, 3, 3, 3))
    q, aux = O.b4step2(prm, q, pt.aux, 1.0, 1.0)
    dt = 0.02
    cfl, fm, fp, gm, gp = O.step2(prm, q, aux, 1.0, 1.0, dt)
    I = slice(2, -2)
    dq = -dt * (fm[0, 3:-1, I] - fp[0, I, I]) - dt * (gm[0, I, 3:-1] - gp[0, I, I])
    assert abs(dq.sum()) <= 1e-13 * q[0, I, I].sum()
    assert 0 < cfl < 1


def test_step_transpose_symmetry(O):
    """The y-sweep of a field equals the x-sweep of its transpose with
    hu <-> hv and taudir_x <-> taudir_y (summation order differs: 1e-13)."""
    pt = slope_patch(24, 24, slope=0.4, phi_deg=25.0, seed=5)
    prm = O.params_from(pt.params)
    q = O.bc_fill(pt.q, (1, 1, 1, 1))
    q, aux, dt, _, _ = O.run(prm, q, pt.aux, 1.0, 1.0, 5, 1e-3, 0.9, 1e9)
    q = O.bc_fill(q, (1, 1, 1, 1))
    qt = np.asfortranarray(q.transpose(0, 2, 1)[[0, 2, 1, 3, 4, 5, 6]])
    at = np.asfortranarray(aux.transpose(0, 2, 1)[[0, 1, 2, 3, 5, 4]])
    a, _, ca = O.stepgrid(O.params_from(pt.params), q, aux, 1.0, 1.0, dt)
    b, _, cb = O.stepgrid(O.params_from(pt.params), qt, at, 1.0, 1.0, dt)
    bt = np.asfortranarray(b.transpose(0, 2, 1)[[0, 2, 1, 3, 4, 5, 6]])
    scale = np.abs(a).max(axis=(1, 2))[:, None, None] + 1e-300
    assert np.max(np.abs(a - bt) / scale) < 1e-12
    assert ca == pytest.approx(cb, rel=1e-14)


def test_contact_wave_stays_still(O):
    """Reference test A (tests/eigenstructure_check/check.py:41-83): an
    m-jump in still water on a flat bed, order 1, transverse 0, walls left
    and at y, extrap right, 2 s: max|u| <= 1e-9, h spread <= 1e-9, interface
    does not move."""
    p = config0_params()
    p.phi_deg, p.kref, p.m_crit, p.delta, p.mu = 35.0, 5e-9, 0.64, 0.001, 0.005
    p.order, p.transverse = 1, 0
    mx, my, dx = 200, 10, 0.05
    x = -5.0 + (np.arange(mx) + 0.5) * dx
    m = np.where(x < 0.0, 0.55, 0.60)[:, None] * np.ones((1, my))
    pt = W.make_patch(np.zeros((mx, my)), np.ones((mx, my)), m, 1.0, p, dx, dx)
    prm = O.params_from(p)
    bc = (3, 1, 3, 3)
    q, aux, dt, cfl, dts = O.run(prm, pt.q, pt.aux, dx, dx, 400, 1e-5, 0.4, 0.01, bc)
    assert dts.sum() > 2.0 - 1e-9 or len(dts) == 400
    h = q[0, 2:-2, 2:-2]
    u = q[1, 2:-2, 2:-2] / h
    assert np.abs(u).max() <= 1e-9
    assert h.max() - h.min() <= 1e-9
    mm = q[3, 2:-2, 5] / q[0, 2:-2, 5]
    assert x[np.argmin(np.abs(mm - 0.575))] == x[np.argmin(np.abs(m[:, 0] - 0.575))]


def test_dam_front_bound(O):
    """Reference test B (check.py:86-121): a dry dam break (h = 2 m behind
    x = 5 m, phi = 0) never outruns 5 + 2 sqrt(2 g) t + 0.1 m."""
    p = config0_params()
    p.phi_deg, p.kref, p.delta = 0.0, 5e-9, 0.001
    p.order, p.transverse = 1, 0
    mx, my, dx = 400, 4, 0.05
    x = (np.arange(mx) + 0.5) * dx
    h = np.where(x < 5.0, 2.0, 0.0)[:, None] * np.ones((1, my))
    pt = W.make_patch(np.zeros((mx, my)), h, 0.55, 1.0, p, dx, dx)
    prm = O.params_from(p)
    q = pt.q
    t, dt = 0.0, 1e-5
    for _ in range(8):
        q, aux, dt_next, cfl, dts = O.run(prm, q, pt.aux, dx, dx, 20, dt, 0.4, 0.01, (3, 1, 3, 3))
        t += dts.sum()
        dt = dt_next
        wet = q[0, 2:-2, 2] > 1e-3
        front = x[wet].max()
        assert front <= 5.0 + 2 * math.sqrt(9.81 * 2.0) * t + 0.1
    assert front > 5.5          # and it does move


def test_bc_fill_rules(O):
    """Ghost cells: extrap copies the nearest interior cell; wall mirrors it
    with the normal momentum negated (bc2amr)."""
    rng = np.random.default_rng(4)
    q = np.asfortranarray(rng.standard_normal((7, 9 + 4, 6 + 4)))
    e = O.bc_fill(q, (1, 1, 1, 1))
    I = slice(2, -2)
    assert np.array_equal(e[:, 0, I], q[:, 2, I]) and np.array_equal(e[:, 1, I], q[:, 2, I])
    assert np.array_equal(e[:, -1, I], q[:, -3, I]) and np.array_equal(e[:, I, 0], e[:, I, 2])
    w = O.bc_fill(q, (3, 3, 3, 3))
    assert np.array_equal(w[0, 1, I], q[0, 2, I]) and np.array_equal(w[0, 0, I], q[0, 3, I])
    assert np.array_equal(w[1, 1, I], -q[1, 2, I]) and np.array_equal(w[2, 1, I], q[2, 2, I])
    assert np.array_equal(w[2, I, -2], -w[2, I, -3]) and np.array_equal(w[1, I, -2], w[1, I, -3])
    assert w[1, 0, 0] == q[1, 3, 3] * -1.0 and w[2, 0, 0] == -q[2, 3, 3]


def _ent_state(O, seed=0, vel=1.5, h_e=0.5, bdif=0.0):
    """A moving flow over an erodible layer (aux(7) = h_e), one step's src2."""
    from dclaw_amd import workloads as W
    rng = np.random.default_rng(seed)
    mx, my = 12, 9
    p = W.config0_params()
    h = rng.uniform(0.5, 2.0, (mx, my))
    pt = W.make_patch(np.zeros((mx, my)), h, 0.6, 0.7, p, 1.0, 1.0, maux=7,
                      h_e=np.full((mx, my), h_e))
    pt.q[1] = vel * pt.q[0] * rng.uniform(0.3, 1.0, pt.q[0].shape)
    pt.q[2] = vel * pt.q[0] * rng.uniform(-0.3, 0.3, pt.q[0].shape)
    pt.q[6] = bdif
    return pt


@pytest.mark.parametrize("h_e", [0.5, 1e-6])
def test_entrainment_exchanges_bed_and_flow(oracle_mod, h_e):
    """src2 with entrainment (method 0) against src2 without, on the same
    state: the flow gains exactly the depth the bed loses (to rounding), the
    erosion never exceeds the layer, hu, hv and hm keep their values from
    before the erosion (the reference does not recompute them,
    src2.f90:255-300), and the solid fraction of the gain is me."""
    O = oracle_mod
    pt = _ent_state(O, h_e=h_e)
    dt = 0.05
    P0 = O.params_from(pt.params)
    q0 = O.src2(P0, pt.q, pt.aux, dt)
    pe = pt.params
    pe.entrainment, pe.entrainment_method, pe.entrainment_rate = 1, 0, 0.4
    q1 = O.src2(O.params_from(pe), pt.q, pt.aux, dt)
    I = (slice(2, -2), slice(2, -2))
    dh = q1[0][I] - q0[0][I]
    db = q1[6][I] - q0[6][I]
    assert (db > 0).all()
    np.testing.assert_allclose(dh, db, rtol=0, atol=8 * np.spacing(q1[0][I]).max())
    assert (q1[6][I] <= h_e * (1 + 1e-15)).all()
    assert np.array_equal(q1[1], q0[1]) and np.array_equal(q1[2], q0[2])
    assert np.array_equal(q1[3], q0[3])
    assert (q1[5][I] > q0[5][I]).all()                  # hchi gains dh*chie
    np.testing.assert_allclose(q1[5][I] - q0[5][I], db * pe.chie, rtol=1e-9, atol=1e-15)


def test_entrainment_threshold_and_method(oracle_mod):
    """No erosion below the 0.1 m/s threshold or with no layer left, and
    entrainment_method = 1 changes nothing (src2.f90:258, entrainment.f90:54-57)."""
    O = oracle_mod
    dt = 0.05
    for kw in (dict(vel=0.05), dict(h_e=0.3, bdif=0.3), dict()):
        pt = _ent_state(O, **kw)
        q0 = O.src2(O.params_from(pt.params), pt.q, pt.aux, dt)
        pe = pt.params
        pe.entrainment = 1
        pe.entrainment_method = 0 if kw else 1
        q1 = O.src2(O.params_from(pe), pt.q, pt.aux, dt)
        assert np.array_equal(q0, q1), kw
