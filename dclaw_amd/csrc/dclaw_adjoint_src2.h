// dclaw_adjoint_src2.h -- the reverse (vector-Jacobian product) of the src2
// cell body written out, for the adjoint step's src2 stage, and the reverses
// of qfix, qfix_cmass and setvars it is built from (the prepare stage's
// reverse, prepare_vjp in dclaw_adjoint.h, uses them too).
//
// src2_cell (dclaw_physics_body.inc; src2.f90:100-300) with the relaxation
// source (mp_update_relax_Dclaw4, mp_update.f90:710-789; src2method -1, 0,
// 1), without entrainment erosion, Manning friction or curvature (those
// configurations keep the seeded products).  Given the adjoints ly of the
// cell's six outputs it returns the adjoints lx of the six inputs and the
// adjoints of every parameter field the body reads (ParAdj), from one
// recomputation of the forward values and one backward sweep.  The
// derivative conventions are the tangent-linear type's (dclaw_tangent.h):
// at max/min the branch the value takes, sqrt'(0) = 0, and so on, so the
// result equals the seeded product vjp<7, 7>(Src2F) up to rounding.
#pragma once

#include "dclaw_physics.h"

namespace dclaw_adj {

// Adjoints of the parameter fields the cell physics reads.
struct ParAdj {
    double grav, rho_s, rho_f, mu, phi, kref, mref, delta_kr_order, alpha_c, sigma_0, delta,
        m_crit, c1, speed_limit;
};

DC_HD void pa_zero(ParAdj& a)
{
    a.grav = a.rho_s = a.rho_f = a.mu = a.phi = a.kref = a.mref = a.delta_kr_order = 0.0;
    a.alpha_c = a.sigma_0 = a.delta = a.m_crit = a.c1 = a.speed_limit = 0.0;
}

// ParAdj . dP: the directional derivative along a parameter direction.
DC_HD double pa_dot(const ParAdj& a, const dclaw_params& d)
{
    double s = a.grav * d.grav;
    s = s + a.rho_s * d.rho_s;
    s = s + a.rho_f * d.rho_f;
    s = s + a.mu * d.mu;
    s = s + a.phi * d.phi;
    s = s + a.kref * d.kref;
    s = s + a.mref * d.mref;
    s = s + a.delta_kr_order * d.delta_kr_order;
    s = s + a.alpha_c * d.alpha_c;
    s = s + a.sigma_0 * d.sigma_0;
    s = s + a.delta * d.delta;
    s = s + a.m_crit * d.m_crit;
    s = s + a.c1 * d.c1;
    s = s + a.speed_limit * d.speed_limit;
    return s;
}

// ------------------------------------------------------------------ setvars
struct SvFwd {
    double r1, rho, tpow, kr, ex, kperm, s_pre, sig_eff, sig_0, alphainv, vn, shear, sd, Nden,
        Nnum, sN, sM, m_eq, w, T, tau0, th, tau;
    bool seg, am1, meq_const, shear_on;
};

// setvars (digclaw_module.f90:394-491) with the intermediates its reverse
// reads; the values are setvars' own, operation for operation.
DC_HD void setvars_fwd(const dclaw_params& P, double h, double u, double v, double m, double p,
                       double chi, double gz, SvFwd& s)
{
    s.r1 = P.rho_s - P.rho_f;
    s.rho = m * s.r1 + P.rho_f;
    s.kr = 1.0;
    s.tpow = 0.0;
    s.seg = P.segregation == 1;
    if (s.seg) {
        s.tpow = P.delta_kr_order * 2.0 * (chi - 0.5);
        s.kr = ::dcm_pow(10.0, s.tpow);
    }
    s.ex = ::dcm_exp(-((m - P.mref) / 0.04));
    s.kperm = s.kr * P.kref * s.ex;
    s.s_pre = s.rho * gz * h - p;
    s.sig_eff = 0.0 > s.s_pre ? 0.0 : s.s_pre;
    s.am1 = P.alphamethod == 1;
    s.sig_0 = s.am1 ? 0.50 * P.alpha_c * P.rho_f * (P.rho_s - P.rho_f) * gz * h / s.rho
                    : P.sigma_0;
    s.alphainv = m * (s.sig_eff + s.sig_0) / P.alpha_c;
    s.vn = sqrt(u * u + v * v);
    s.shear_on = h > P.dry_tolerance;
    s.shear = s.shear_on ? 2.0 * s.vn / h : 0.0;
    s.sd = s.shear * P.delta;
    s.Nden = P.rho_s * (s.sd * s.sd) + s.sig_eff;
    s.Nnum = P.mu * s.shear;
    s.meq_const = s.Nnum <= 0.0;
    s.sN = s.sM = 0.0;
    if (s.meq_const) {
        s.m_eq = P.m_crit;
    } else {
        s.sN = sqrt(s.Nden);
        s.sM = sqrt(s.Nnum);
        s.m_eq = P.m_crit * (s.sN / (s.sN + s.sM));
    }
    s.w = P.c1 * (m - s.m_eq);
    s.T = ::dcm_tan(P.phi + ::dcm_atan(s.w));
    s.tau0 = s.sig_eff * s.T;
    s.th = ::dcm_tanh(100.0 * (m - 0.1));
    s.tau = s.tau0 * 0.5 * (1.0 + s.th);
}

// Adjoints b_* of setvars' outputs (rho, kperm, alphainv, m_eq, tanpsi, tau)
// added into the adjoints of its inputs (bh .. bchi, bgz) and of the
// parameters.
DC_HD void setvars_rev(const dclaw_params& P, double h, double u, double v, double m, double chi,
                       double gz, const SvFwd& s, double b_rho, double b_kperm, double b_ai,
                       double b_meq, double b_tanpsi, double b_tau, double& bh, double& bu,
                       double& bv, double& bm, double& bp, double& bchi, double& bgz,
                       ParAdj& pa)
{
    // tau = (tau0 0.5) (1 + tanh(100 (m - 0.1)))
    const double b_tau0 = b_tau * 0.5 * (1.0 + s.th);
    bm += b_tau * (s.tau0 * 0.5) * (1.0 - s.th * s.th) * 100.0;
    // tau0 = sig_eff tan(phi + atan(w))
    double b_sig = b_tau0 * s.T;
    const double b_arg = b_tau0 * s.sig_eff * (1.0 + s.T * s.T);
    pa.phi += b_arg;
    const double b_w = b_arg / (1.0 + s.w * s.w) + b_tanpsi;
    // w = c1 (m - m_eq)
    pa.c1 += b_w * (m - s.m_eq);
    bm += b_w * P.c1;
    const double bme = b_meq - b_w * P.c1;
    double b_Nden = 0.0, b_Nnum = 0.0;
    if (s.meq_const) {
        pa.m_crit += bme;
    } else {
        const double S = s.sN + s.sM, rS = 1.0 / S;
        pa.m_crit += bme * (s.sN * rS);
        const double bf = bme * P.m_crit;
        const double b_sN = bf * s.sM * (rS * rS), b_sM = -bf * s.sN * (rS * rS);
        if (s.sN > 0.0) b_Nden += b_sN * (0.5 / s.sN);
        if (s.sM > 0.0) b_Nnum += b_sM * (0.5 / s.sM);
    }
    // Nnum = mu shear; Nden = rho_s sd^2 + sig_eff; sd = shear delta
    pa.mu += b_Nnum * s.shear;
    double b_shear = b_Nnum * P.mu;
    pa.rho_s += b_Nden * (s.sd * s.sd);
    const double b_sd = b_Nden * P.rho_s * 2.0 * s.sd;
    b_sig += b_Nden;
    b_shear += b_sd * P.delta;
    pa.delta += b_sd * s.shear;
    if (s.shear_on) {
        const double rh = 1.0 / h;
        const double b_vn = b_shear * 2.0 * rh;
        bh -= b_shear * s.shear * rh;
        if (s.vn > 0.0) {
            const double rvn = 1.0 / s.vn;
            bu += b_vn * u * rvn;
            bv += b_vn * v * rvn;
        }
    }
    // alphainv = m (sig_eff + sig_0) / alpha_c
    const double rac = 1.0 / P.alpha_c;
    bm += b_ai * (s.sig_eff + s.sig_0) * rac;
    const double b_S = b_ai * m * rac;
    pa.alpha_c -= b_ai * s.alphainv * rac;
    b_sig += b_S;
    double b_r = b_rho;
    if (s.am1) {
        // sig_0 = 0.5 alpha_c rho_f (rho_s - rho_f) gz h / rho
        const double rrho = 1.0 / s.rho;
        const double ghr = gz * h * rrho;
        const double c0 = 0.5 * P.alpha_c * P.rho_f * s.r1;
        pa.alpha_c += b_S * 0.5 * P.rho_f * s.r1 * ghr;
        pa.rho_f += b_S * 0.5 * P.alpha_c * (s.r1 - P.rho_f) * ghr;
        pa.rho_s += b_S * 0.5 * P.alpha_c * P.rho_f * ghr;
        bgz += b_S * c0 * h * rrho;
        bh += b_S * c0 * gz * rrho;
        b_r -= b_S * s.sig_0 * rrho;
    } else {
        pa.sigma_0 += b_S;
    }
    // sig_eff = max(0, rho gz h - p)
    if (!(0.0 > s.s_pre)) {
        b_r += b_sig * gz * h;
        bgz += b_sig * s.rho * h;
        bh += b_sig * s.rho * gz;
        bp -= b_sig;
    }
    // kperm = kr kref exp(-(m - mref) / 0.04)
    pa.kref += b_kperm * s.kr * s.ex;
    const double b_a = b_kperm * s.kr * P.kref * s.ex;
    bm -= b_a * 25.0;   // d/dm of -(m - mref) / 0.04
    pa.mref += b_a * 25.0;
    if (s.seg) {
        const double b_t = b_kperm * P.kref * s.ex * s.kr * ::dcm_log(10.0);
        pa.delta_kr_order += b_t * 2.0 * (chi - 0.5);
        bchi += b_t * P.delta_kr_order * 2.0;
    }
    // rho = m (rho_s - rho_f) + rho_f
    bm += b_r * s.r1;
    pa.rho_s += b_r * m;
    pa.rho_f += b_r * (1.0 - m);
}

// --------------------------------------------------------------------- qfix
struct QfFwd {
    double h, hu, hv, p_in, u, v, m0, m, chi0, chi, rho, pmax, p1, p;
    bool dry;
};

// qfix (digclaw_module.f90:285-345): outputs h, hu, hv (unchanged), hm = h m,
// p, hchi = h chi, u, v, m, chi, rho.
DC_HD void qfix_fwd(const dclaw_params& P, double h, double hu, double hv, double hm, double p,
                    double hchi, double gz, QfFwd& f)
{
    f.h = h; f.hu = hu; f.hv = hv; f.p_in = p;
    f.dry = h <= P.dry_tolerance;
    if (f.dry) {
        f.u = f.v = f.m0 = f.m = f.chi0 = f.chi = f.rho = f.pmax = f.p1 = f.p = 0.0;
        return;
    }
    f.u = hu / h;
    f.v = hv / h;
    f.m0 = hm / h;
    f.chi0 = hchi / h;
    double m = f.m0 > 0.0 ? f.m0 : 0.0;
    f.m = m < 1.0 ? m : 1.0;
    double c = f.chi0 > 0.0 ? f.chi0 : 0.0;
    f.chi = c < 1.0 ? c : 1.0;
    f.rho = P.rho_s * f.m + (1.0 - f.m) * P.rho_f;
    f.pmax = f.rho * gz * h;
    f.p1 = 0.0 > p ? 0.0 : p;
    f.p = f.pmax < f.p1 ? f.pmax : f.p1;
}

// Output adjoints in, input adjoints out (in place: bh, bhu, bhv become the
// adjoints of the inputs h, hu, hv; bhm, bp, bhchi of hm, p, hchi).
DC_HD void qfix_rev(const dclaw_params& P, const QfFwd& f, double gz, double& bh, double& bhu,
                    double& bhv, double& bhm, double& bp, double& bhchi, double bu, double bv,
                    double bm, double bchi, double brho, double& bgz, ParAdj& pa)
{
    if (f.dry) {
        bh = bhu = bhv = bhm = bp = bhchi = 0.0;
        return;
    }
    double bpmax = 0.0, bp1 = 0.0;
    if (f.pmax < f.p1) bpmax = bp;
    else bp1 = bp;
    const double bp_in = (0.0 > f.p_in) ? 0.0 : bp1;
    const double br = brho + bpmax * gz * f.h;
    bgz += bpmax * f.rho * f.h;
    bh += bpmax * f.rho * gz;
    double bmt = bm + br * (P.rho_s - P.rho_f);
    pa.rho_s += br * f.m;
    pa.rho_f += br * (1.0 - f.m);
    bh += bhchi * f.chi;
    const double bct = bchi + bhchi * f.h;
    const double bchi0 = (f.chi0 > 0.0 && f.chi0 < 1.0) ? bct : 0.0;
    bh += bhm * f.m;
    bmt += bhm * f.h;
    const double bm0 = (f.m0 > 0.0 && f.m0 < 1.0) ? bmt : 0.0;
    const double rh = 1.0 / f.h;   // one division (an FP64 division is 11 instructions)
    bhm = bm0 * rh;
    bh -= bm0 * f.m0 * rh;
    bhchi = bchi0 * rh;
    bh -= bchi0 * f.chi0 * rh;
    bhu += bu * rh;
    bh -= bu * f.u * rh;
    bhv += bv * rh;
    bh -= bv * f.v * rh;
    bp = bp_in;
}

// --------------------------------------------------------------- qfix_cmass
struct QcFwd {
    double m_in, m1, rho, h, pemax, pemin, pe_in, pe1, pe, p, u, v;
};

// qfix_cmass (digclaw_module.f90:353-387) from its live inputs m, p_exc, u,
// v, rhoh, gz; outputs h, m, rho, p_exc, p, hu = h u, hv = h v, hm = h m.
DC_HD void qcm_fwd(const dclaw_params& P, double m, double p_exc, double u, double v,
                   double rhoh, double gz, QcFwd& c)
{
    c.m_in = m;
    c.pe_in = p_exc;
    c.u = u;
    c.v = v;
    double m1 = m > 0.0 ? m : 0.0;
    c.m1 = m1 < 1.0 ? m1 : 1.0;
    c.rho = c.m1 * (P.rho_s - P.rho_f) + P.rho_f;
    c.h = rhoh / c.rho;
    c.pemax = rhoh * gz - P.rho_f * c.h * gz;
    c.pemin = 0.0 - P.rho_f * c.h * gz;
    c.pe1 = c.pemin > p_exc ? c.pemin : p_exc;
    c.pe = c.pemax < c.pe1 ? c.pemax : c.pe1;
    c.p = P.rho_f * gz * c.h + c.pe;
}

DC_HD void qcm_rev(const dclaw_params& P, const QcFwd& c, double rhoh, double gz, double bh,
                   double bm, double bp, double brho, double bpe, double bhu, double bhv,
                   double bhm, double& bm_in, double& bpe_in, double& bu, double& bv,
                   double& brhoh, double& bgz, ParAdj& pa)
{
    double bht = bh + bhm * c.m1 + bhv * c.v + bhu * c.u;
    double bm1 = bm + bhm * c.h;
    bu += bhu * c.h;
    bv += bhv * c.h;
    const double bpt = bpe + bp;
    pa.rho_f += bp * gz * c.h;
    bgz += bp * P.rho_f * c.h;
    bht += bp * P.rho_f * gz;
    double bpemax = 0.0, bpe1 = 0.0;
    if (c.pemax < c.pe1) bpemax = bpt;
    else bpe1 = bpt;
    double bpemin = 0.0;
    bpe_in = 0.0;
    if (c.pemin > c.pe_in) bpemin = bpe1;
    else bpe_in = bpe1;
    pa.rho_f -= bpemin * c.h * gz;
    bht -= bpemin * P.rho_f * gz;
    bgz -= bpemin * P.rho_f * c.h;
    brhoh += bpemax * gz;
    bgz += bpemax * (rhoh - P.rho_f * c.h);
    pa.rho_f -= bpemax * c.h * gz;
    bht -= bpemax * P.rho_f * gz;
    const double rrho = 1.0 / c.rho;
    brhoh += bht * rrho;
    const double brt = brho - bht * c.h * rrho;
    bm1 += brt * (P.rho_s - P.rho_f);
    pa.rho_s += brt * c.m1;
    pa.rho_f += brt * (1.0 - c.m1);
    bm_in = (c.m_in > 0.0 && c.m_in < 1.0) ? bm1 : 0.0;
}

// ---------------------------------------------------------------------- src2
// Whether src2_vjp covers this configuration.
DC_HD bool src2_vjp_covers(const dclaw_params& P)
{
    if (P.src2method < -1 || P.src2method > 1) return false;
    if (P.entrainment == 1 && P.entrainment_method == 0) return false;
    if (P.friction_forcing && P.manning_coefficient > 0.0) return false;
    return P.curvature == 0;
}

// src2's Coulomb friction (src2.f90:160-176) on the qfix'd state: the
// forward values its reverse reads.
struct FrFwd {
    double hvn0, vn_a, pre, vn_b, ea, E0, vn_c, hvn, hu2, hv2, u2, v2;
    bool fr;
};

DC_HD void friction_fwd(const dclaw_params& P, double dt, double h0, double hu0, double hv0,
                        double m1, double u1, double v1, double tau1, double rhoh, FrFwd& f)
{
    f.hvn0 = sqrt(hu0 * hu0 + hv0 * hv0);
    f.fr = f.hvn0 > 0.0;
    f.vn_a = f.pre = f.vn_b = f.ea = f.E0 = f.vn_c = f.hvn = 0.0;
    f.hu2 = hu0; f.hv2 = hv0; f.u2 = u1; f.v2 = v1;
    if (f.fr) {
        f.vn_a = f.hvn0 / h0;
        f.pre = f.vn_a - dt * tau1 / rhoh;
        f.vn_b = 0.0 > f.pre ? 0.0 : f.pre;
        f.ea = -((1.0 - m1) * 2.0 * P.mu * dt / (h0 * rhoh));
        f.E0 = ::dcm_exp(f.ea);
        f.vn_c = f.vn_b * f.E0;
        f.hvn = h0 * f.vn_c;
        f.hu2 = f.hvn * hu0 / f.hvn0;
        f.hv2 = f.hvn * hv0 / f.hvn0;
        f.u2 = f.hu2 / h0;
        f.v2 = f.hv2 / h0;
    }
}

// x: the cell's state (h > dry_tolerance) before src2; ly: adjoints of the
// six outputs h, hu, hv, hm, pb, hchi; lx: adjoints of the six inputs (bdif
// is the caller's: it passes through); pa: parameter adjoints (added).
DC_HD void src2_vjp(const dclaw_params& P, double dt, const double* x, const double* ly,
                    double* lx, ParAdj& pa)
{
    const double gz = P.grav;
    const double h0 = x[0], hu0 = x[1], hv0 = x[2], hm0 = x[3], p0 = x[4], hchi0 = x[5];
    const double rhoh = hm0 * P.rho_s + (h0 - hm0) * P.rho_f;
    if (hm0 < 0.0) {
        // src2.f90:113: the cell is left as it is
        for (int k = 0; k < 6; k++) lx[k] = ly[k];
        return;
    }
    // ---- forward values.  The first qfix and setvars are evaluated again
    // for their reverses at the end (fewer values held across the whole
    // reverse: registers, not scratch, at two waves per SIMD)
    double u1, v1, m1, chi1, p1, tau1;
    {
        QfFwd q1;
        qfix_fwd(P, h0, hu0, hv0, hm0, p0, hchi0, gz, q1);
        u1 = q1.u; v1 = q1.v; m1 = q1.m; chi1 = q1.chi; p1 = q1.p;
        SvFwd s1;
        setvars_fwd(P, h0, u1, v1, m1, p1, chi1, gz, s1);
        tau1 = s1.tau;
    }
    // friction (evaluated again for its reverse at the end)
    double u2, v2;
    bool fr;
    {
        FrFwd f;
        friction_fwd(P, dt, h0, hu0, hv0, m1, u1, v1, tau1, rhoh, f);
        u2 = f.u2;
        v2 = f.v2;
        fr = f.fr;
    }
    // the reverse divides through reciprocals (one division each; an FP64
    // division is a chain of 11 instructions); the forward values above keep
    // src2's own divisions, since they select its branches
    const double rh0 = 1.0 / h0, rrhoh = 1.0 / rhoh;
    double b_h = 0.0, b_u2 = 0.0, b_v2 = 0.0, b_m1 = 0.0, b_p1 = 0.0, b_chi1 = 0.0;
    double b_rhoh = 0.0, b_gz = 0.0;
    if (P.src2method == -1) {
        // outputs h0, h0 u2, h0 v2, h0 m1, p1, h0 chi1
        b_h += ly[0] + ly[1] * u2 + ly[2] * v2 + ly[3] * m1 + ly[5] * chi1;
        b_u2 += ly[1] * h0;
        b_v2 += ly[2] * h0;
        b_m1 += ly[3] * h0;
        b_chi1 += ly[5] * h0;
        b_p1 += ly[4];
    } else {
        // ---- mp_update_relax forward
        double s2w, s2ai, s2kp;   // the second setvars' tanpsi, alphainv, kperm
        {
            SvFwd s2;
            setvars_fwd(P, h0, u2, v2, m1, p1, chi1, gz, s2);
            s2w = s2.w; s2ai = s2.alphainv; s2kp = s2.kperm;
        }
        const double hu3 = h0 * u2, hv3 = h0 * v2, hm3 = h0 * m1, hchi3 = h0 * chi1;
        const double sq3 = sqrt(hu3 * hu3 + hv3 * hv3);
        const double vn3 = sq3 / h0;
        const double c3 = dt * 3.0 * vn3 * s2w * s2ai;
        const double p3 = p1 - c3 / h0;
        double u4, v4, m4, chi4, rho4, p4;
        {
            QfFwd q2;
            qfix_fwd(P, h0, hu3, hv3, hm3, p3, hchi3, gz, q2);
            u4 = q2.u; v4 = q2.v; m4 = q2.m; chi4 = q2.chi; rho4 = q2.rho; p4 = q2.p;
        }
        const double t2n = (rho4 - P.rho_f) * P.rho_f * gz;
        const double zeta = 3.0 * s2ai / (h0 * 2.0) + t2n / (4.0 * rho4);
        const bool mu_on = P.mu > 1e-16;
        const double muc = mu_on ? P.mu : 1e-16;
        const double kr1 = -(zeta * 2.0 * s2kp / (h0 * muc));
        const double p_eq = P.rho_f * h0 * gz;
        const double E1 = ::dcm_exp(kr1 * dt);
        const double p5 = p_eq + (p4 - p_eq) * E1;
        const double pexc = p5 - P.rho_f * gz * h0;
        QcFwd qc;
        qcm_fwd(P, m4, pexc, u4, v4, rhoh, gz, qc);
        const double h6 = qc.h, hu6 = h6 * u4, hv6 = h6 * v4, hm6 = h6 * qc.m1;
        const double g6 = s2kp / (P.mu * h6);
        const double D = -(2.0 * g6 * qc.pe);
        const double kr2 = D * (qc.rho - P.rho_f) / rhoh;
        const double e2 = -(dt * D * P.rho_f / rhoh);
        const double E2 = ::dcm_exp(e2);
        const double hm7 = hm6 * E2;
        double h7, hu7, hv7, hchi7, E3 = 0.0;
        if (P.src2method == 0) {
            h7 = h6 + h6 * kr2 * dt;
            hchi7 = h7 * chi4;
            E3 = ::dcm_exp(dt * kr2);
            hu7 = hu6 * E3;
            hv7 = hv6 * E3;
        } else {
            h7 = (rhoh - hm7 * (P.rho_s - P.rho_f)) / P.rho_f;
            hu7 = h7 * u4;
            hv7 = h7 * v4;
            hchi7 = h7 * chi4;
        }
        QfFwd q3;
        qfix_fwd(P, h7, hu7, hv7, hm7, qc.p, hchi7, gz, q3);
        if (q3.dry) {
            // src2 zeroes the cell: no dependence on the inputs
            for (int k = 0; k < 6; k++) lx[k] = 0.0;
            return;
        }
        const double rrf = 1.0 / P.rho_f, rrho4 = 1.0 / rho4, rh6 = 1.0 / h6;
        const double rmu = 1.0 / P.mu, rmuc = 1.0 / muc;
        // ---- reverse of src2's outputs h7, h7 uF, h7 vF, h7 mF, pF, h7 chiF
        double bh7 = ly[0] + ly[1] * q3.u + ly[2] * q3.v + ly[3] * q3.m + ly[5] * q3.chi;
        double bhu7 = 0.0, bhv7 = 0.0, bhm7 = 0.0, bp6 = ly[4], bhchi7 = 0.0;
        qfix_rev(P, q3, gz, bh7, bhu7, bhv7, bhm7, bp6, bhchi7, ly[1] * h7, ly[2] * h7,
                 ly[3] * h7, ly[5] * h7, 0.0, b_gz, pa);
        double b_h6 = 0.0, b_hu6 = 0.0, b_hv6 = 0.0, b_kr2 = 0.0, b_chi4 = 0.0;
        double b_u4 = 0.0, b_v4 = 0.0, b_hm7 = bhm7;
        if (P.src2method == 0) {
            b_hu6 += bhu7 * E3;
            b_hv6 += bhv7 * E3;
            const double bE3 = bhu7 * hu6 + bhv7 * hv6;
            b_kr2 += bE3 * E3 * dt;
            bh7 += bhchi7 * chi4;
            b_chi4 += bhchi7 * h7;
            b_h6 += bh7 * (1.0 + kr2 * dt);
            b_kr2 += bh7 * h6 * dt;
        } else {
            bh7 += bhchi7 * chi4 + bhu7 * u4 + bhv7 * v4;
            b_chi4 += bhchi7 * h7;
            b_u4 += bhu7 * h7;
            b_v4 += bhv7 * h7;
            b_rhoh += bh7 * rrf;
            b_hm7 -= bh7 * (P.rho_s - P.rho_f) * rrf;
            pa.rho_s -= bh7 * hm7 * rrf;
            pa.rho_f += bh7 * (hm7 - h7) * rrf;
        }
        // hm7 = hm6 E2, E2 = exp(-(dt D rho_f / rhoh))
        const double b_hm6 = b_hm7 * E2;
        const double b_e2 = b_hm7 * hm6 * E2;
        double b_D = -b_e2 * dt * P.rho_f * rrhoh;
        pa.rho_f -= b_e2 * dt * D * rrhoh;
        b_rhoh -= b_e2 * e2 * rrhoh;
        // kr2 = D (rho6 - rho_f) / rhoh
        b_D += b_kr2 * (qc.rho - P.rho_f) * rrhoh;
        const double b_rho6 = b_kr2 * D * rrhoh;
        pa.rho_f -= b_kr2 * D * rrhoh;
        b_rhoh -= b_kr2 * kr2 * rrhoh;
        // D = -2 g6 pe6, g6 = kperm / (mu h6)
        const double b_pe6 = -2.0 * b_D * g6;
        const double b_g6 = -2.0 * b_D * qc.pe;
        double b_kperm = b_g6 * (rmu * rh6);
        pa.mu -= b_g6 * g6 * rmu;
        b_h6 -= b_g6 * g6 * rh6;
        // qfix_cmass
        double b_m4 = 0.0, b_pexc = 0.0;
        qcm_rev(P, qc, rhoh, gz, b_h6, 0.0, bp6, b_rho6, b_pe6, b_hu6, b_hv6, b_hm6, b_m4,
                b_pexc, b_u4, b_v4, b_rhoh, b_gz, pa);
        // pexc = p5 - rho_f gz h0; p5 = p_eq + (p4 - p_eq) E1
        pa.rho_f -= b_pexc * gz * h0;
        b_gz -= b_pexc * P.rho_f * h0;
        b_h -= b_pexc * P.rho_f * gz;
        const double b_peq = b_pexc * (1.0 - E1);
        double b_p4 = b_pexc * E1;
        const double b_kr1 = b_pexc * (p4 - p_eq) * E1 * dt;
        pa.rho_f += b_peq * h0 * gz;
        b_h += b_peq * P.rho_f * gz;
        b_gz += b_peq * P.rho_f * h0;
        // kr1 = -(zeta 2 kperm / (h0 muc))
        const double b_zeta = -b_kr1 * 2.0 * s2kp * (rh0 * rmuc);
        b_kperm -= b_kr1 * zeta * 2.0 * (rh0 * rmuc);
        b_h -= b_kr1 * kr1 * rh0;
        if (mu_on) pa.mu -= b_kr1 * kr1 * rmuc;
        // zeta = 3 alphainv / (2 h0) + (rho4 - rho_f) rho_f gz / (4 rho4)
        double b_ai = b_zeta * 1.5 * rh0;
        b_h -= b_zeta * (1.5 * s2ai * rh0) * rh0;
        const double q4 = 0.25 * rrho4;
        const double t2 = t2n * q4;
        const double b_rho4 = b_zeta * (P.rho_f * gz * q4 - t2 * rrho4);
        pa.rho_f += b_zeta * (rho4 - 2.0 * P.rho_f) * gz * q4;
        b_gz += b_zeta * (rho4 - P.rho_f) * P.rho_f * q4;
        // qfix of mp_update (its h, hu, hv, hm, hchi outputs are not read later)
        double bq2h = 0.0, bq2hu = 0.0, bq2hv = 0.0, bq2hm = 0.0, bq2p = b_p4, bq2hchi = 0.0;
        QfFwd q2;   // evaluated again (the forward kept only its outputs)
        qfix_fwd(P, h0, hu3, hv3, hm3, p3, hchi3, gz, q2);
        qfix_rev(P, q2, gz, bq2h, bq2hu, bq2hv, bq2hm, bq2p, bq2hchi, b_u4, b_v4, b_m4, b_chi4,
                 b_rho4, b_gz, pa);
        b_h += bq2h;
        // p3 = p1 - (dt 3 vn3 tanpsi alphainv) / h0
        b_p1 += bq2p;
        const double b_c3 = -bq2p * rh0;
        b_h += bq2p * c3 * (rh0 * rh0);
        const double b_vn3 = b_c3 * dt * 3.0 * s2w * s2ai;
        const double b_tanpsi = b_c3 * dt * 3.0 * vn3 * s2ai;
        b_ai += b_c3 * dt * 3.0 * vn3 * s2w;
        double b_hu3 = bq2hu, b_hv3 = bq2hv;
        const double b_hm3 = bq2hm, b_hchi3 = bq2hchi;
        b_h -= b_vn3 * vn3 * rh0;
        if (sq3 > 0.0) {
            const double rsq3 = 1.0 / sq3;
            b_hu3 += b_vn3 * rh0 * hu3 * rsq3;
            b_hv3 += b_vn3 * rh0 * hv3 * rsq3;
        }
        b_h += b_hu3 * u2 + b_hv3 * v2 + b_hm3 * m1 + b_hchi3 * chi1;
        b_u2 += b_hu3 * h0;
        b_v2 += b_hv3 * h0;
        b_m1 += b_hm3 * h0;
        b_chi1 += b_hchi3 * h0;
        SvFwd s2;   // evaluated again
        setvars_fwd(P, h0, u2, v2, m1, p1, chi1, gz, s2);
        setvars_rev(P, h0, u2, v2, m1, chi1, gz, s2, 0.0, b_kperm, b_ai, 0.0, b_tanpsi, 0.0,
                    b_h, b_u2, b_v2, b_m1, b_p1, b_chi1, b_gz, pa);
    }
    // ---- friction (src2.f90:160-176)
    double b_hu0 = 0.0, b_hv0 = 0.0, b_u1 = 0.0, b_v1 = 0.0, b_tau1 = 0.0;
    if (fr) {
        FrFwd f;
        friction_fwd(P, dt, h0, hu0, hv0, m1, u1, v1, tau1, rhoh, f);
        const double hvn0 = f.hvn0, vn_a = f.vn_a, pre = f.pre, vn_b = f.vn_b, ea = f.ea;
        const double E0 = f.E0, vn_c = f.vn_c, hvn = f.hvn, hu2 = f.hu2, hv2 = f.hv2;
        const double rhv0 = 1.0 / hvn0;
        const double b_hu2 = b_u2 * rh0, b_hv2 = b_v2 * rh0;
        b_h -= b_u2 * u2 * rh0 + b_v2 * v2 * rh0;
        const double b_hvn = b_hu2 * hu0 * rhv0 + b_hv2 * hv0 * rhv0;
        b_hu0 += b_hu2 * hvn * rhv0;
        b_hv0 += b_hv2 * hvn * rhv0;
        double b_hvn0 = -(b_hu2 * hu2 + b_hv2 * hv2) * rhv0;
        b_h += b_hvn * vn_c;
        const double b_vnc = b_hvn * h0;
        const double b_vnb = b_vnc * E0;
        const double b_ea = b_vnc * vn_b * E0;
        b_m1 += b_ea * 2.0 * P.mu * dt * (rh0 * rrhoh);
        pa.mu -= b_ea * (1.0 - m1) * 2.0 * dt * (rh0 * rrhoh);
        b_h -= b_ea * ea * rh0;
        b_rhoh -= b_ea * ea * rrhoh;
        if (!(0.0 > pre)) {
            b_hvn0 += b_vnb * rh0;
            b_h -= b_vnb * vn_a * rh0;
            b_tau1 -= b_vnb * dt * rrhoh;
            b_rhoh += b_vnb * dt * tau1 * (rrhoh * rrhoh);
        }
        b_hu0 += b_hvn0 * hu0 * rhv0;
        b_hv0 += b_hvn0 * hv0 * rhv0;
    } else {
        b_u1 += b_u2;
        b_v1 += b_v2;
    }
    SvFwd s1;
    setvars_fwd(P, h0, u1, v1, m1, p1, chi1, gz, s1);
    setvars_rev(P, h0, u1, v1, m1, chi1, gz, s1, 0.0, 0.0, 0.0, 0.0, 0.0, b_tau1, b_h, b_u1,
                b_v1, b_m1, b_p1, b_chi1, b_gz, pa);
    // ---- the first qfix (its hm, hchi, rho outputs are not read later)
    double bhm = 0.0, bhchi = 0.0, bp = b_p1;
    QfFwd q1;
    qfix_fwd(P, h0, hu0, hv0, hm0, p0, hchi0, gz, q1);
    qfix_rev(P, q1, gz, b_h, b_hu0, b_hv0, bhm, bp, bhchi, b_u1, b_v1, b_m1, b_chi1, 0.0, b_gz,
             pa);
    // ---- rhoh = hm0 rho_s + (h0 - hm0) rho_f
    lx[0] = b_h + b_rhoh * P.rho_f;
    lx[1] = b_hu0;
    lx[2] = b_hv0;
    lx[3] = bhm + b_rhoh * (P.rho_s - P.rho_f);
    lx[4] = bp;
    lx[5] = bhchi;
    pa.rho_s += b_rhoh * hm0;
    pa.rho_f += b_rhoh * (h0 - hm0);
    pa.grav += b_gz;
}

}  // namespace dclaw_adj
