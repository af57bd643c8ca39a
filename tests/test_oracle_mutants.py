"""Mutation check of the oracle's pins (CPU only).

Each mutant is a copy of oracle/dclaw_oracle.c with one plausible
transcription slip (a dropped factor, a halved coefficient, a flipped test, a
swapped operand) in a line the GPU shares with the oracle only through bit
parity.  The copy is compiled like the oracle (gcc, -ffp-contract=off) into a
temporary directory and loaded through oracle.library(); the test passes if
at least one pin -- the physics pins of test_oracle_physics_pins.py or the
pins of test_oracle_pins.py -- rejects the mutant.  A mutant that every pin
accepts would be a slip the parity chain cannot see.
"""
import os
import subprocess

import pytest

import test_oracle_physics_pins as PP
import test_oracle_pins as OP

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "dclaw_oracle.c")

# (name, text in the oracle, replacement)
MUTANTS = [
    ("flux2 order-2 correction (1-|s|dtdx) -> (1-|s|dtdx/2)",
     "(1.0 - fabs(sv) * dtdx1d[i - 1 + mbc - 1]) * fv",
     "(1.0 - 0.5 * fabs(sv) * dtdx1d[i - 1 + mbc - 1]) * fv"),
    ("flux2 order-2 correction without the limiter",
     "if (limit) oracle_limiter(P, mx, mbc, fwave, s);", ""),
    ("flux2 transverse 2 folds cqxx with the wrong sign",
     "S1(amdq, MEQN, m, i) = S1(amdq, MEQN, m, i) + S1(cqxx, MEQN, m, i);",
     "S1(amdq, MEQN, m, i) = S1(amdq, MEQN, m, i) - S1(cqxx, MEQN, m, i);"),
    ("rpt2 pb row of wave 1: gamma rho g -> rho g",
     "r[4][1] = m; r[5][1] = gamma * rho * g;", "r[4][1] = m; r[5][1] = rho * g;"),
    ("rpt2 pb row of wave 3: gamma rho g -> rho g",
     "r[4][3] = m; r[5][3] = gamma * rho * g;", "r[4][3] = m; r[5][3] = rho * g;"),
    ("rpt2 hm row of wave 3: m -> m/2",
     "r[4][3] = m; r[5][3]", "r[4][3] = 0.5 * m; r[5][3]"),
    ("rpt2 transverse speeds v -+ sqrt(gh) -> v -+ sqrt(gh/2)",
     "s[1] = v - sqrt(g * h);", "s[1] = v - sqrt(0.5 * g * h);"),
    ("rpt2 state of the wrong cell for imp",
     "int ic = (imp == 1) ? i - 1 : i;", "int ic = (imp == 1) ? i : i - 1;"),
    ("solver pb f-wave of wave 1 without gamma_L",
     "FW(5, 1) = FW(1, 1) * gammaL * rhoL * grav * cos(theta1);",
     "FW(5, 1) = FW(1, 1) * rhoL * grav * cos(theta1);"),
    ("solver revert fold direction flipped",
     "if (hustar_interface <= 0.0) {", "if (hustar_interface >= 0.0) {"),
    ("solver moving-bed source without s1s2tilde/s1s2bar",
     "else source2dx = -(gz * hbar * delb * s1s2tilde / s1s2bar);",
     "else source2dx = -(gz * hbar * delb);"),
    ("solver hm f-wave of wave 3 with the left m",
     "FW(4, 3) = FW(1, 3) * mR;", "FW(4, 3) = FW(1, 3) * mL;"),
    ("solver hv f-wave of wave 1 with the right v",
     "FW(3, 1) = FW(1, 1) * vL;", "FW(3, 1) = FW(1, 1) * vR;"),
    ("rpn2 hu f-wave of the middle wave dropped",
     "FWV(mhu, 2) = F6(2, 2);", "FWV(mhu, 2) = 0.0;"),
    ("rpn2 speed of wave 5 = middle speed",
     "SS(5) = sw[2];", "SS(5) = sw[1];"),
    ("setvars m_eq with Nnum and Nden swapped",
     "*m_eq = P->m_crit * (sqrt(Nden) / (sqrt(Nden) + sqrt(Nnum)));",
     "*m_eq = P->m_crit * (sqrt(Nnum) / (sqrt(Nden) + sqrt(Nnum)));"),
    ("setvars permeability exponent sign",
     "*kperm = kr_chi * P->kref * D_EXP(-((m - P->mref) / 0.04));",
     "*kperm = kr_chi * P->kref * D_EXP(((m - P->mref) / 0.04));"),
    ("setvars friction tan(phi + psi) -> tan(phi)",
     "*tau = sig_eff * D_TAN(P->phi + psi);", "*tau = sig_eff * D_TAN(P->phi);"),
    ("setvars shear rate 2|v|/h -> |v|/h",
     "if (h > P->dry_tolerance) shear = 2.0 * vnorm / h;",
     "if (h > P->dry_tolerance) shear = vnorm / h;"),
    ("src2 Coulomb decrement without rho (tau/h)",
     "vnorm = dmax(0.0, vnorm - dt * tau / rhoh);", "vnorm = dmax(0.0, vnorm - dt * tau / h);"),
    ("src2 viscous decay without (1 - m)",
     "vnorm = vnorm * D_EXP(-((1.0 - m) * 2.0 * mu * dt / (h * rhoh)));",
     "vnorm = vnorm * D_EXP(-(2.0 * mu * dt / (h * rhoh)));"),
    ("relaxation rate zeta: 3/alpha -> 2/alpha",
     "double zeta = 3.0 * alphainv / (*h * 2.0)", "double zeta = 2.0 * alphainv / (*h * 2.0)"),
    ("relaxation rate zeta: rho_f g (rho - rho_f)/(4 rho) term dropped",
     "+ (rho - rho_f) * rho_f * gz / (4.0 * rho);", ";"),
    ("dilatancy rate 3|v| -> 1.5|v|",
     "*p = *p - dt * 3.0 * vnorm * tanpsi * alphainv / *h;",
     "*p = *p - dt * 1.5 * vnorm * tanpsi * alphainv / *h;"),
    ("hm update with rho_s for rho_f",
     "hm = hm * D_EXP(-(dt * D * rho_f / (rhoh)));", "hm = hm * D_EXP(-(dt * D * rho_s / (rhoh)));"),
    ("depth update krate without (rho - rho_f)",
     "krate = D * (rho - rho_f) / rhoh;", "krate = D * rho / rhoh;"),
    ("speed limit rescales hu only",
     "P2(q, MEQN, 3, i, j) = P2(q, MEQN, 3, i, j) * sratio;", ""),
    ("qfix_cmass lower bound halved",
     "double p_exc_min = 0.0 - P->rho_f * *h * gz;",
     "double p_exc_min = 0.0 - 0.5 * P->rho_f * *h * gz;"),
    ("static branch with half the friction",
     "} else if (0.5 * fabs(taudirR * tauR / rhoR + tauL * taudirR / rhoL) > fabs(del[2] - source2dx)) {",
     "} else if (0.25 * fabs(taudirR * tauR / rhoR + tauL * taudirR / rhoL) > fabs(del[2] - source2dx)) {"),
    ("calc_taudir driving force from the right gradient only",
     "if (FxL * FxR > 0.0) Fx = dsign(dmin(fabs(FxL), fabs(FxR)), FxL);",
     "if (FxL * FxR > 0.0) Fx = FxR;"),
    ("calc_taudir still cell: taudir_x scaled by dy",
     "P2(aux, maux, I_TAUDIR_X, i, j) = -(dx * FxL / sqrt(FxL * FxL + Fy * Fy));",
     "P2(aux, maux, I_TAUDIR_X, i, j) = -(dy * FxL / sqrt(FxL * FxL + Fy * Fy));"),
    ("setvars effective stress from rho_f",
     "double sig_eff = dmax(0.0, *rho * gz * h - p);",
     "double sig_eff = dmax(0.0, P->rho_f * gz * h - p);"),
    ("FE_4quad kp0 with the sign of its second term flipped",
     "(3.0 * alphainv * rho0 / rhoh - 1.5 * rho_f * gz * h0 * (rho0 - rho_f) / rhoh);",
     "(3.0 * alphainv * rho0 / rhoh + 1.5 * rho_f * gz * h0 * (rho0 - rho_f) / rhoh);"),
    ("FE_4quad dilatancy forcing 3|v| -> 1.5|v|",
     "const double c_d0 = -(3.0 * vnorm", "const double c_d0 = -(1.5 * vnorm"),
    ("FE_4quad m rate with rho0 for rho0^2",
     "const double km0 = ((2.0 * kperm * (rho0 * rho0)) / (mu * (rhoh * rhoh)));",
     "const double km0 = ((2.0 * kperm * (rho0)) / (mu * (rhoh * rhoh)));"),
    ("curvature gravity reset for every cell",
     "                gz = gz + gacc;", "                gz = P->grav + gacc;"),
    ("curvature acceleration with u instead of u^2",
     "dmax(uc * uc * b_xx + vc * vc * b_yy", "dmax(uc * b_xx + vc * vc * b_yy"),
    ("qfix pb bound rho g h -> rho_f g h",
     "double pmax = *rho * gz * *h;", "double pmax = P->rho_f * gz * *h;"),
]

# cheap checks first; the convergence studies last
PIN_FUNCS = list(PP.PINS[:-2]) + [
    OP.test_riemanntype_known_answer, OP.test_solver_fwave_sums_flat_bed,
    OP.test_solver_lake_at_rest, OP.test_solver_static_branch,
    OP.test_rpt2_reproduces_transverse_jacobian, OP.test_qfix_admissibility,
    OP.test_setvars_special_values, OP.test_src2_water_viscous_decay,
    OP.test_src2_coulomb_stop_and_direction, OP.test_src2_pressure_relaxes_toward_hydrostatic,
    OP.test_step2_conserves_mass_with_walls, OP.test_step_transpose_symmetry,
] + list(PP.PINS[-2:])


def _build(tmp_path, old, new) -> str:
    src = open(SRC).read()
    assert src.count(old) == 1, f"mutation site not unique: {old!r}"
    d = tmp_path / "mut"
    d.mkdir()
    c = d / "dclaw_oracle.c"
    c.write_text(src.replace(old, new))
    for h in ("dclaw_oracle.h",):
        (d / h).write_text(open(os.path.join(ROOT, "oracle", h)).read())
    so = d / "libmut.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-std=c11",
                    "-I" + os.path.join(ROOT, "oracle"), "-shared", "-o", str(so), str(c), "-lm"],
                   check=True, cwd=os.path.join(ROOT, "oracle"))
    return str(so)


def _first_rejection(O):
    for f in PIN_FUNCS:
        try:
            f(O)
        except (AssertionError, FloatingPointError, ZeroDivisionError, ValueError) as e:
            return f.__name__, str(e).splitlines()[0][:100] if str(e) else ""
    return None


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutant_is_rejected(oracle_mod, tmp_path, name, old, new):
    so = _build(tmp_path, old, new)
    with oracle_mod.library(so):
        rej = _first_rejection(oracle_mod)
    assert rej is not None, f"no pin rejects the mutant: {name}"
    print(f"{name}: rejected by {rej[0]}")


def test_unmutated_oracle_passes_every_pin(oracle_mod):
    assert _first_rejection(oracle_mod) is None
