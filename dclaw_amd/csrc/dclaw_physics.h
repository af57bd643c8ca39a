// dclaw_physics.h -- point and edge physics of the D-Claw step, FP64.
//
// Every routine is __host__ __device__ so the same code runs in the gfx950
// kernels and in the host-side lane emulator used by the tests.  The
// floating-point evaluation order follows the reference Fortran statement by
// statement (left-to-right products and sums, x**2 as x*x, unary minus
// outside the product); build with -ffp-contract=off so that no FMA
// contraction changes a rounding.  MAX/MIN are (a > b ? a : b) and
// (a < b ? a : b); SIGN(a, b) is copysign(|a|, b).
//
// Structurally zero f-wave components are not carried: with
// delta_huvm_revert = .true. (riemannsolvers_dclaw.f:66) the middle solver
// wave keeps only its h and normal-momentum parts (:513-516), so rpn2's
// waves 3 and 4 are identically zero (rpn2_dclaw.f:303-315) and wave 2 has
// only h and normal momentum (:296-301); no wave has a bdif component.
// Dropping +-0 terms from sums that start at +0 leaves every sum bit-exact.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "dclaw_hip.h"

#define DCM_FN __host__ __device__ __forceinline__
#include "dcmath.h"

namespace dclaw {

#define DC_HD __host__ __device__ __forceinline__

DC_HD double dmax(double a, double b) { return a > b ? a : b; }
DC_HD double dmin(double a, double b) { return a < b ? a : b; }
DC_HD double dsign(double a, double b) { return copysign(fabs(a), b); }

// Conserved state of one cell, components in q-index order (bdif apart).
struct Cell {
    double h, hu, hv, hm, p, hchi;
};

// qfix: digclaw_module.f90:285-345.  Updates the conserved state in place
// and returns the primitives u = hu/h, v = hv/h, m, chi, rho.
DC_HD void qfix(const dclaw_params& P, double& h, double& hu, double& hv, double& hm, double& p,
                double& hchi, double& u, double& v, double& m, double& chi, double& rho,
                double gz)
{
    if (h <= P.dry_tolerance) {
        h = 0.0; hu = 0.0; hv = 0.0; hm = 0.0; p = 0.0; hchi = 0.0;
        u = 0.0; v = 0.0; m = 0.0; chi = 0.0; rho = 0.0;
        return;
    }
    u = hu / h;
    v = hv / h;
    m = hm / h;
    chi = hchi / h;
    m = dmax(m, 0.0);
    m = dmin(m, 1.0);
    hm = h * m;
    chi = dmax(chi, 0.0);
    chi = dmin(chi, 1.0);
    hchi = h * chi;
    rho = P.rho_s * m + (1.0 - m) * P.rho_f;
    const double pmax = rho * gz * h;
    p = dmax(0.0, p);
    p = dmin(pmax, p);
}

// qfix_cmass: digclaw_module.f90:353-387
DC_HD void qfix_cmass(const dclaw_params& P, double& h, double& m, double& p, double& rho,
                      double& p_exc, double& hu, double& hv, double& hm, double u, double v,
                      double rhoh, double gz)
{
    m = dmax(m, 0.0);
    m = dmin(m, 1.0);
    rho = m * (P.rho_s - P.rho_f) + P.rho_f;
    h = rhoh / rho;
    const double p_exc_max = rhoh * gz - P.rho_f * h * gz;
    const double p_exc_min = 0.0 - P.rho_f * h * gz;
    p_exc = dmax(p_exc_min, p_exc);
    p_exc = dmin(p_exc_max, p_exc);
    p = P.rho_f * gz * h + p_exc;
    hu = h * u;
    hv = h * v;
    hm = h * m;
}

struct Vars {
    double rho, kperm, alphainv, m_eq, tanpsi, tau;
};

// setvars: digclaw_module.f90:394-491.  P.phi plays the module global phi
// that calc_taudir sets from aux(i_phi) (:557).
DC_HD Vars setvars(const dclaw_params& P, double h, double u, double v, double m, double p,
                   double chi, double gz)
{
    Vars o;
    o.rho = m * (P.rho_s - P.rho_f) + P.rho_f;
    double kr_chi = 1.0;
    if (P.segregation == 1) kr_chi = dcm_pow(10.0, P.delta_kr_order * 2.0 * (chi - 0.5));
    o.kperm = kr_chi * P.kref * dcm_exp(-((m - P.mref) / 0.04));
    const double sig_eff = dmax(0.0, o.rho * gz * h - p);
    double sig_0 = P.sigma_0;
    if (P.alphamethod == 1)
        sig_0 = 0.50 * P.alpha_c * P.rho_f * (P.rho_s - P.rho_f) * gz * h / o.rho;
    o.alphainv = m * (sig_eff + sig_0) / P.alpha_c;
    const double vnorm = sqrt(u * u + v * v);
    double shear;
    if (h > P.dry_tolerance) shear = 2.0 * vnorm / h;
    else shear = 0.0;
    const double sd = shear * P.delta;
    const double Nden = P.rho_s * (sd * sd) + sig_eff;
    const double Nnum = P.mu * shear;
    if (Nnum <= 0.0) o.m_eq = P.m_crit;
    else o.m_eq = P.m_crit * (sqrt(Nden) / (sqrt(Nden) + sqrt(Nnum)));
    o.tanpsi = P.c1 * (m - o.m_eq);
    const double psi = dcm_atan(P.c1 * (m - o.m_eq));
    double tau = sig_eff * dcm_tan(P.phi + psi);
    o.tau = tau * 0.5 * (1.0 + dcm_tanh(100.0 * (m - 0.1)));
    return o;
}

struct RType {
    double hm, s1m, s2m;
    bool rare1, rare2;
};

// riemanntype: riemannsolvers_dclaw.f:526-637 (maxiter = 1 at every call)
DC_HD RType riemanntype(double hL, double hR, double uL, double uR, double drytol, double g)
{
    RType o;
    const double h_min = dmin(hR, hL);
    const double h_max = dmax(hR, hL);
    const double delu = uR - uL;
    if (h_min <= drytol) {
        o.hm = 0.0;
        o.s1m = uR + uL - 2.0 * sqrt(g * hR) + 2.0 * sqrt(g * hL);
        o.s2m = uR + uL - 2.0 * sqrt(g * hR) + 2.0 * sqrt(g * hL);
        if (hL <= 0.0) { o.rare2 = true; o.rare1 = false; }
        else { o.rare1 = true; o.rare2 = false; }
        return o;
    }
    const double F_min = delu + 2.0 * (sqrt(g * h_min) - sqrt(g * h_max));
    const double F_max = delu + (h_max - h_min) * (sqrt(.5 * g * (h_max + h_min) / (h_max * h_min)));
    if (F_min > 0.0) {
        const double t = dmax(0.0, -delu + 2.0 * (sqrt(g * hL) + sqrt(g * hR)));
        o.hm = (1.0 / (16.0 * g)) * (t * t);
        o.s1m = uL + 2.0 * sqrt(g * hL) - 3.0 * sqrt(g * o.hm);
        o.s2m = uR - 2.0 * sqrt(g * hR) + 3.0 * sqrt(g * o.hm);
        o.rare1 = true;
        o.rare2 = true;
    } else if (F_max <= 0.0) {
        double h0 = h_max;
        {
            const double gL = sqrt(.5 * g * (1.0 / h0 + 1.0 / hL));
            const double gR = sqrt(.5 * g * (1.0 / h0 + 1.0 / hR));
            const double F0 = delu + (h0 - hL) * gL + (h0 - hR) * gR;
            const double dfdh = gL - g * (h0 - hL) / (4.0 * (h0 * h0) * gL) + gR
                                - g * (h0 - hR) / (4.0 * (h0 * h0) * gR);
            const double slope = 2.0 * sqrt(h0) * dfdh;
            const double t = sqrt(h0) - F0 / slope;
            h0 = t * t;
        }
        o.hm = h0;
        const double u1m = uL - (o.hm - hL) * sqrt((.5 * g) * (1.0 / o.hm + 1.0 / hL));
        const double u2m = uR + (o.hm - hR) * sqrt((.5 * g) * (1.0 / o.hm + 1.0 / hR));
        o.s1m = u1m - sqrt(g * o.hm);
        o.s2m = u2m + sqrt(g * o.hm);
        o.rare1 = false;
        o.rare2 = false;
    } else {
        double h0 = h_min;
        {
            const double F0 = delu + 2.0 * (sqrt(g * h0) - sqrt(g * h_max))
                              + (h0 - h_min) * sqrt(.5 * g * (1.0 / h0 + 1.0 / h_min));
            const double slope = (F_max - F0) / (h_max - h_min);
            h0 = h0 - F0 / slope;
        }
        o.hm = h0;
        if (hL > hR) {
            o.s1m = uL + 2.0 * sqrt(g * hL) - 3.0 * sqrt(g * o.hm);
            o.s2m = uL + 2.0 * sqrt(g * hL) - sqrt(g * o.hm);
            o.rare1 = true;
            o.rare2 = false;
        } else {
            o.s2m = uR - 2.0 * sqrt(g * hR) + 3.0 * sqrt(g * o.hm);
            o.s1m = uR - 2.0 * sqrt(g * hR) + sqrt(g * o.hm);
            o.rare2 = true;
            o.rare1 = false;
        }
    }
    return o;
}

// The waves of one edge as rpn2 stores them (rpn2_dclaw.f:276-315), only the
// structurally nonzero components, each vector in q-index order
// (h, hu, hv, hm, pb, hchi).  s[0] = s(1), s[1] = s(2) = s(3) = s(4),
// s[2] = s(5).  w2h, w2n: wave 2's h and normal-momentum components.
struct EdgeWaves {
    double s[3];
    double w1[6];
    double w5[6];
    double w2h, w2n;
};

// Inputs of one normal Riemann problem after rpn2's qfix calls: conserved
// state (q-index order) and primitives (u normal, v transverse velocity).
struct Side {
    double h, hu, hv, hm, p, hchi;  // q-index order (hu = x momentum)
    double u, v, m, chi;             // qfix primitives, u = normal / h
    double b;                        // aux(1) - bdif
    double taudir, fs;               // aux(i_taudir_{x|y}), aux(i_fsphi)
    double rho, tau;                 // setvars(h, u, v, m, p, chi) of this state; read
                                     // only when h > dry_tolerance
};

// The two setvars outputs rpn2 and calc_taudir read.  setvars depends on the
// velocity only through u*u + v*v, so the value is the same for the x and
// y sweeps and for a wall-mirrored copy of the state (u -> -u): one
// evaluation per cell state serves every edge that reads it.
struct RhoTau {
    double rho, tau;
};

DC_HD RhoTau rho_tau(const dclaw_params& P, double h, double u, double v, double m, double p,
                     double chi)
{
    const Vars s = setvars(P, h, u, v, m, p, chi, P.grav);
    RhoTau o;
    o.rho = s.rho;
    o.tau = s.tau;
    return o;
}

// riemann_dig2_aug_sswave_ez (riemannsolvers_dclaw.f:3-521) fused with the
// rpn2 per-edge body (rpn2_dclaw.f:125-315): dry skip, wall/inundation tests,
// solver, wall masking and the 3 -> 5 wave mapping.  ixy = 1 for an x-edge
// (normal momentum hu), 2 for a y-edge (normal momentum hv).
DC_HD EdgeWaves rpn2_edge(const dclaw_params& P, int ixy, const Side& SL, const Side& SR)
{
    EdgeWaves W;
    W.s[0] = W.s[1] = W.s[2] = 0.0;
    for (int k = 0; k < 6; k++) { W.w1[k] = 0.0; W.w5[k] = 0.0; }
    W.w2h = 0.0;
    W.w2n = 0.0;
    const double drytol = P.dry_tolerance;
    const double grav = P.grav;
    if (SL.h <= drytol && SR.h <= drytol) return W;   // rpn2_dclaw.f:125

    const bool xd = (ixy == 1);
    double hL = SL.h, hR = SR.h;
    double huL = xd ? SL.hu : SL.hv, huR = xd ? SR.hu : SR.hv;  // normal
    double hvL = xd ? SL.hv : SL.hu, hvR = xd ? SR.hv : SR.hu;  // transverse
    double hmL = SL.hm, hmR = SR.hm;
    double pL = SL.p, pR = SR.p;
    double bL = SL.b, bR = SR.b;
    double uL = SL.u, uR = SR.u, vL = SL.v, vR = SR.v;
    double mL = SL.m, mR = SR.m, chiL = SL.chi, chiR = SR.chi;
    const double taudirL = SL.taudir, taudirR = SR.taudir;
    const double fsL = SL.fs, fsR = SR.fs;

    // setvars values of the two states as the solver will see them
    double svrL = SL.rho, svtL = SL.tau, svrR = SR.rho, svtR = SR.tau;
    double wall0 = 1.0, wall1 = 1.0, wall2 = 1.0;
    bool wallprob = false;
    bool resetL = false, resetR = false;   // a dry side reset in place (not mirrored)
    if (hR <= drytol) {
        hR = 0.0;
        pR = 0.0;
        mR = mL;
        chiR = chiL;
        resetR = true;
        const RType rt = riemanntype(hL, hL, uL, -uL, drytol, grav);
        const double hstartest = dmax(hL, rt.hm);
        if (hstartest + bL < bR) {
            wall1 = 0.0;
            wall2 = 0.0;
            wallprob = true;
            resetR = false;
            hR = hL; huR = -huL; hvR = hvL; hmR = hmL; bR = bL;
            uR = -uL; vR = vL; mR = mL; pR = pL; chiR = chiL;
            svrR = svrL; svtR = svtL;
        }
    } else if (hL <= drytol) {
        hL = 0.0;
        pL = 0.0;
        mL = mR;
        chiL = chiR;
        resetL = true;
        const RType rt = riemanntype(hR, hR, -uR, uR, drytol, grav);
        const double hstartest = dmax(hR, rt.hm);
        if (hstartest + bR < bL) {
            wall0 = 0.0;
            wall1 = 0.0;
            wallprob = true;
            resetL = false;
            hL = hR; huL = -huR; hvL = hvR; hmL = hmR; mL = mR; bL = bR;
            uL = -uR; vL = vR; pL = pR; chiL = chiR;
            svrL = svrR; svtL = svtR;
        }
    }
    // A reset dry side (h = 0) reaches setvars below only when dry_tolerance
    // <= 0; evaluate it on its reset state then.
    if (drytol <= 0.0 && resetR) {
        const RhoTau t = rho_tau(P, hR, uR, vR, mR, pR, chiR);
        svrR = t.rho; svtR = t.tau;
    }
    if (drytol <= 0.0 && resetL) {
        const RhoTau t = rho_tau(P, hL, uL, vL, mL, pL, chiL);
        svrL = t.rho; svtL = t.tau;
    }

    // ---- riemann_dig2_aug_sswave_ez (thetaL = thetaR = 0) ----
    const double rho_f = P.rho_f, kappa = P.kappa;
    const double criticaltol = dmax(drytol * grav, 1e-6);
    const double criticaltol_2 = sqrt(criticaltol);
    const double gzL = grav * cos(0.0);
    const double gzR = grav * cos(0.0);
    const double gz = grav * cos(0.5 * (0.0 + 0.0));
    const double alpha_seg = 1.0 - P.beta_seg;

    // setvars of each wet side (riemannsolvers_dclaw.f:64-80), evaluated once
    // per cell state by the caller (gzL = gzR = grav: theta = 0)
    (void)gzL;
    (void)gzR;
    double rhoL, rhoR, tauL, tauR, h;
    if (hL >= drytol && hR >= drytol) {
        rhoL = svrL; tauL = svtL; rhoR = svrR; tauR = svtR;
        h = 0.5 * (hL + hR);
    } else if (hL >= drytol) {
        rhoL = svrL; tauL = svtL; rhoR = svrL;
        tauR = 0.5 * tauL;
        h = 0.5 * hL;
    } else {
        rhoL = svrR; rhoR = svrR; tauR = svtR;
        tauL = 0.5 * tauR;
        h = 0.5 * hR;
    }
    tauL = fsL * tauL;
    tauR = fsR * tauR;
    const double rho = 0.5 * (rhoL + rhoR);

    double gamma, gammaL, gammaR;
    if (P.src2method == -1) {
        gamma = rho_f / rho;
        gammaL = rho_f / rhoL;
        gammaR = rho_f / rhoR;
    } else {
        gamma = 0.25 * (rho_f + 3.0 * rho) / rho;
        gammaL = 0.25 * (rho_f + 3.0 * rhoL) / rhoL;
        gammaR = 0.25 * (rho_f + 3.0 * rhoR) / rhoR;
    }
    const double eps = kappa + (1.0 - kappa) * gamma;
    const double geps = gz * eps;

    const double sqhL = sqrt(hL), sqhR = sqrt(hR);
    const double sL = uL - sqrt(geps * hL);
    const double sR = uR + sqrt(geps * hR);
    const double uhat = (sqhL * uL + sqhR * uR) / (sqhR + sqhL);
    const double chat = sqrt(geps * 0.5 * (hR + hL));
    const double sRoe1 = uhat - chat;
    const double sRoe2 = uhat + chat;
    const double sE1 = dmin(sL, sRoe1);
    const double sE2 = dmax(sR, sRoe2);
    const double u = uhat;

    const RType rt = riemanntype(hL, hR, uL, uR, drytol, geps);
    double sw0 = dmin(sE1, rt.s2m);
    double sw2 = dmax(sE2, rt.s1m);
    double sw1 = 0.5 * (sw2 + sw0);

    const double hstarHLL = dmax((huL - huR + sE2 * hR - sE1 * hL) / (sE2 - sE1), 0.0);

    bool rarecorrector = false;
    if (P.rarecorrectortest) {
        const double sdelta = sw2 - sw0;
        double raremin = 0.5;
        const double raremax = 0.9;
        if (rt.rare1 && sE1 * rt.s1m < 0.0) raremin = 0.2;
        if (rt.rare2 && sE2 * rt.s2m < 0.0) raremin = 0.2;
        if (rt.rare1 || rt.rare2) {
            const double rare1st = 3.0 * (sqrt(geps * hL) - sqrt(geps * rt.hm));
            const double rare2st = 3.0 * (sqrt(geps * hR) - sqrt(geps * rt.hm));
            if (dmax(rare1st, rare2st) > raremin * sdelta &&
                dmax(rare1st, rare2st) < raremax * sdelta) {
                rarecorrector = true;
                if (rare1st > rare2st) sw1 = rt.s1m;
                else if (rare2st > rare1st) sw1 = rt.s2m;
                else sw1 = 0.5 * (rt.s1m + rt.s2m);
            }
        }
        if (hstarHLL < dmin(hL, hR) / 5.0) rarecorrector = false;
    }

    const double delb = (bR - bL);
    const double hbar = 0.5 * (hL + hR);
    const double tuu = uL + uR;
    const double s1s2bar = 0.25 * (tuu * tuu) - gz * hbar;
    const double s1s2tilde = dmax(0.0, uL * uR) - gz * hbar;

    bool sonic = false;
    if (fabs(s1s2bar) <= criticaltol) sonic = true;
    else if (s1s2bar * s1s2tilde <= criticaltol * criticaltol) sonic = true;
    else if (s1s2bar * sE1 * sE2 <= criticaltol * criticaltol) sonic = true;
    else if (dmin(fabs(sE1), fabs(sE2)) < criticaltol_2) sonic = true;
    else if (sE1 < criticaltol_2 && rt.s1m > -criticaltol_2) sonic = true;
    else if (sE2 > -criticaltol_2 && rt.s2m < criticaltol_2) sonic = true;
    else if ((uL + sqrt(geps * hL)) * (uR + sqrt(geps * hR)) < 0.0) sonic = true;
    else if ((uL - sqrt(geps * hL)) * (uR - sqrt(geps * hR)) < 0.0) sonic = true;

    double source2dx;
    if (sonic) source2dx = -(gz * hbar * delb);
    else source2dx = -(gz * hbar * delb * s1s2tilde / s1s2bar);
    source2dx = dmin(source2dx, gz * dmax(-(hL * delb), -(hR * delb)));
    source2dx = dmax(source2dx, gz * dmin(-(hL * delb), -(hR * delb)));
    if (fabs(u) <= 0.0) source2dx = -(hbar * gz * delb);

    double deldelh;
    if (sonic) deldelh = -delb;
    else deldelh = delb * gz * hbar / s1s2bar;
    if (sE1 < -criticaltol && sE2 > criticaltol) {
        deldelh = dmin(deldelh, hstarHLL * (sE2 - sE1) / sE2);
        deldelh = dmax(deldelh, hstarHLL * (sE2 - sE1) / sE1);
    } else if (sE1 >= criticaltol) {
        deldelh = dmin(deldelh, hstarHLL * (sE2 - sE1) / sE1);
        deldelh = dmax(deldelh, -hL);
    } else if (sE2 <= -criticaltol) {
        deldelh = dmin(deldelh, hR);
        deldelh = dmax(deldelh, hstarHLL * (sE2 - sE1) / sE2);
    }

    double del0 = hR - hL - deldelh;
    double del1 = huR - huL;
    double del2 = hR * (uR * uR) + 0.5 * kappa * gz * (hR * hR)
                  - (hL * (uL * uL) + 0.5 * kappa * gz * (hL * hL));
    del2 = del2 + (1.0 - kappa) * h * (pR - pL) / rho;
    double del4 = -(gamma * rho * gz * u * (hR - hL)) + gamma * rho * gz * del1 + u * (pR - pL);

    const double vnorm = sqrt(uR * uR + uL * uL + vR * vR + vL * vL);
    bool is_static = false;
    double tausource;
    if (vnorm > 0.0) {
        tausource = 0.0;
    } else if (0.5 * fabs(taudirR * tauR / rhoR + tauL * taudirR / rhoL) > fabs(del2 - source2dx)) {
        tausource = del2 - source2dx;
        del1 = 0.0;
        del0 = 0.0;
        del4 = 0.0;
        is_static = true;
    } else {
        tausource = 0.5 * ((taudirR * tauR / rhoR) + (tauL * taudirR / rhoL));
        tausource = dsign(tausource, del2 - source2dx);
    }
    if (wallprob) tausource = 0.0;
    del2 = del2 - source2dx - tausource;

    // theta1 = theta3 = 0 in every branch (bed_normal = 0): cos(0) = 1.
    const double a = sw0, b = sw1, c = sw2;
    double beta1, beta2, beta3, R12, R22;
    if (!rarecorrector) {
        beta1 = (c * del0 - del1) / (c - a);
        beta2 = a * c * del0 - (a + c) * del1 + del2;
        beta3 = (del1 - a * del0) / (c - a);
        R12 = 0.0;
        R22 = 1.0;
    } else {
        beta1 = (b * c * del0 - (b + c) * del1 + del2) / (a * a - a * b - a * c + b * c);
        beta2 = (-(a * c * del0) + (a + c) * del1 - del2) / (a * b - a * c - b * b + b * c);
        beta3 = (a * b * del0 - (a + b) * del1 + del2) / (a * b - a * c - b * c + c * c);
        R12 = sw1;
        R22 = sw1 * sw1;
    }
    // fw(m, mw) in solver order: 1 h, 2 normal, 3 transverse, 4 hm, 5 pb, 6 hchi
    double f11 = beta1 * a, f21 = beta1 * (a * a);
    double f12 = beta2 * R12, f22 = beta2 * R22;
    double f13 = beta3 * c, f23 = beta3 * (c * c);

    double f41 = f11 * mL;
    double f43 = f13 * mR;
    const double f42 = hmR * uR - hmL * uL - f41 - f43;
    double f31 = f11 * vL;
    double f33 = f13 * vR;
    const double f32 = hvR * uR - hvL * uL - f31 - f33;
    double f51 = f11 * gammaL * rhoL * grav * cos(0.0);
    double f53 = f13 * gammaR * rhoR * grav * cos(0.0);
    double f52 = del4 - f53 - f51;
    const double seg_L = chiL * hL * uL * (1.0 + (1.0 - alpha_seg) * (1.0 - chiL));
    const double seg_R = chiR * hR * uR * (1.0 + (1.0 - alpha_seg) * (1.0 - chiR));
    double f61 = f11 * chiL * (1.0 + (1.0 - alpha_seg) * (1.0 - chiL));
    double f63 = f13 * chiR * (1.0 + (1.0 - alpha_seg) * (1.0 - chiR));
    const double f62 = seg_R - seg_L - f61 - f63;

    // delta_huvm_revert (riemannsolvers_dclaw.f:471-517)
    const double hustar_interface = huL + f11;
    if (!is_static) {
        const double u_interface = hustar_interface / hstarHLL;
        del4 = -(gamma * rho * gz * u_interface * (hR - hL)) + gamma * rho * gz * del1
               + u_interface * (pR - pL);
        f52 = del4 - f53 - f51;
    }
    if (hustar_interface <= 0.0) {
        f31 = f31 + f32;
        f41 = f41 + f42;
        f51 = f51 + f52;
        f61 = f61 + f62;
    } else {
        f33 = f33 + f32;
        f43 = f43 + f42;
        f53 = f53 + f52;
        f63 = f63 + f62;
    }

    // wall masking (rpn2_dclaw.f:267-272) and mapping to q-index order
    W.s[0] = sw0 * wall0;
    W.s[1] = sw1 * wall1;
    W.s[2] = sw2 * wall2;
    const double n1 = f21 * wall0, t1 = f31 * wall0;
    const double n3 = f23 * wall2, t3 = f33 * wall2;
    W.w1[0] = f11 * wall0;
    W.w1[1] = xd ? n1 : t1;
    W.w1[2] = xd ? t1 : n1;
    W.w1[3] = f41 * wall0;
    W.w1[4] = f51 * wall0;
    W.w1[5] = f61 * wall0;
    W.w5[0] = f13 * wall2;
    W.w5[1] = xd ? n3 : t3;
    W.w5[2] = xd ? t3 : n3;
    W.w5[3] = f43 * wall2;
    W.w5[4] = f53 * wall2;
    W.w5[5] = f63 * wall2;
    W.w2h = f12 * wall1;
    W.w2n = f22 * wall1;
    return W;
}

// Limiter function of the AMRClaw inline limiter (Clawpack 5 inlinelimiter.f,
// called at flux2fw.f:188), for a wave with wnorm2 != 0 and ratio r.
DC_HD double limiter_phi(int mthlim, double r)
{
    switch (mthlim) {
    case 1: return dmax(0.0, dmin(1.0, r));
    case 2: return dmax(dmax(0.0, dmin(1.0, 2.0 * r)), dmin(2.0, r));
    case 3: return (r + fabs(r)) / (1.0 + fabs(r));
    case 4: {
        const double c = (1.0 + r) / 2.0;
        return dmax(0.0, dmin(dmin(c, 2.0), 2.0 * r));
    }
    default: return r;
    }
}

// Transverse splitting in one cell (rpt2_dclaw.f:78-260) for the two
// fluctuations that use this cell's Jacobian: imp = 1 for the left-going
// amdq of the edge to its right, imp = 2 for the right-going apdq of the
// edge to its left.  Returns G(k, side) = (0 - g_imp1) - g_imp2 with
// g = 0.5*dtdx*b{m|p}asdq, i.e. flux2's gaddm/gaddp accumulation for this
// cell (flux2fw.f:256-282); side 1 = bmasdq ("below"), 2 = bpasdq ("above").
// Cell state is the sweep's post-rpn2 state; topo1/topo3 are aux(1) of the
// neighbouring slices.  Vectors are q-index order (h, hu, hv, hm, pb, hchi).
struct Trans {
    double g1[6], g2[6];
};

DC_HD void rpt2_cell_split(const dclaw_params& P, int ixy, double h, double hu, double hv,
                           double hm, double hchi, double b, double bdif, double topo1,
                           double topo3, const double* asdq, bool active, double* bm,
                           double* bp)
{
    for (int k = 0; k < 6; k++) { bm[k] = 0.0; bp[k] = 0.0; }
    if (!active || h <= P.dry_tolerance) return;
    const bool xd = (ixy == 1);
    const double u = (xd ? hu : hv) / h;
    const double v = (xd ? hv : hu) / h;
    const double m = hm / h;
    const double chi = hchi / h;
    const double eta = h + b - bdif;
    const double g = P.grav;
    const double rho = m * P.rho_s + (1.0 - m) * P.rho_f;
    double gamma;
    if (P.src2method == -1) gamma = P.rho_f / rho;
    else gamma = 0.25 * (P.rho_f + 3.0 * rho) / rho;
    if (eta < dmin(topo1, topo3)) return;
    const double sq = sqrt(g * h);
    const double s1 = v - sq, s2 = v, s3 = v + sq;
    const int mu = xd ? 1 : 2, mv = xd ? 2 : 1;   // 0-based q-index of normal / transverse
    const double delf1 = asdq[0], delf2 = asdq[mu], delf3 = asdq[mv];
    const double beta1 = (s3 * delf1 - delf3) / (s3 - s1);
    const double beta2 = -(u * delf1) + delf2;
    const double beta3 = (delf3 - s1 * delf1) / (s3 - s1);
    const double r5 = gamma * rho * g;
    // eigenvectors (q-index): r1 = (1, u|s1, s1|u, m, r5, chi) etc.
    double r1[6], r2[6], r3[6];
    r1[0] = 1.0; r1[mu] = u; r1[mv] = s1; r1[3] = m; r1[4] = r5; r1[5] = chi;
    r3[0] = 1.0; r3[mu] = u; r3[mv] = s3; r3[3] = m; r3[4] = r5; r3[5] = chi;
    r2[0] = 0.0; r2[mu] = 1.0; r2[mv] = 0.0; r2[3] = 0.0; r2[4] = 0.0; r2[5] = 0.0;
    const double s[3] = {s1, s2, s3};
    const double be[3] = {beta1, beta2, beta3};
    const double* rr[3] = {r1, r2, r3};
    for (int mw = 0; mw < 3; mw++) {
        if (s[mw] < 0.0 && eta >= topo1) {
            for (int k = 0; k < 6; k++) bm[k] = bm[k] + 1.0 * s[mw] * be[mw] * rr[mw][k];
        } else if (s[mw] > 0.0 && eta >= topo3) {
            for (int k = 0; k < 6; k++) bp[k] = bp[k] + 1.0 * s[mw] * be[mw] * rr[mw][k];
        }
    }
}

// mp_update_relax_Dclaw4: mp_update.f90:710-789
DC_HD void mp_update_relax(const dclaw_params& P, double dt, double& h, double& u, double& v,
                           double& m, double& p, double& chi, double rhoh, double gz)
{
    const double rho_f = P.rho_f, rho_s = P.rho_s, mu = P.mu;
    Vars sv = setvars(P, h, u, v, m, p, chi, gz);
    double rho = sv.rho;
    const double kperm = sv.kperm, alphainv = sv.alphainv, tanpsi = sv.tanpsi;
    double hu = h * u;
    double hv = h * v;
    double hm = h * m;
    double hchi = h * chi;
    const double vnorm = sqrt(hu * hu + hv * hv) / h;
    p = p - dt * 3.0 * vnorm * tanpsi * alphainv / h;
    qfix(P, h, hu, hv, hm, p, hchi, u, v, m, chi, rho, gz);

    const double zeta = 3.0 * alphainv / (h * 2.0) + (rho - rho_f) * rho_f * gz / (4.0 * rho);
    double krate = -(zeta * 2.0 * kperm / (h * dmax(mu, 1e-16)));
    const double p_eq = rho_f * h * gz;
    p = p_eq + (p - p_eq) * dcm_exp(krate * dt);

    double p_exc = p - rho_f * gz * h;
    qfix_cmass(P, h, m, p, rho, p_exc, hu, hv, hm, u, v, rhoh, gz);
    const double D = -(2.0 * (kperm / (mu * h)) * p_exc);
    krate = D * (rho - rho_f) / rhoh;
    hm = hm * dcm_exp(-(dt * D * rho_f / (rhoh)));

    if (P.src2method == 0) {
        h = h + h * krate * dt;
        hchi = h * chi;
        hu = hu * dcm_exp(dt * krate);
        hv = hv * dcm_exp(dt * krate);
        qfix(P, h, hu, hv, hm, p, hchi, u, v, m, chi, rho, gz);
    } else if (P.src2method == 1) {
        h = (rhoh - hm * (rho_s - rho_f)) / rho_f;
        hu = h * u;
        hv = h * v;
        hchi = h * chi;
        qfix(P, h, hu, hv, hm, p, hchi, u, v, m, chi, rho, gz);
    }
}

// ent_dclaw4 (entrainment.f90:27-100): erosion of the bed layer into the
// flow at one cell.  Its b_x, b_y arguments are not used by the reference, so
// they are not passed here.
DC_HD void ent_dclaw4(const dclaw_params& P, double dt, double& h, double u, double v, double& m,
                      double& p, double& rho, double& hchi, double gz, double tau,
                      double& b_eroded, double b_remaining)
{
    double coeff;
    if (P.friction_forcing) coeff = P.manning_coefficient;
    else coeff = 0.0;
    const double vnorm = sqrt(u * u + v * v);
    const double vlow = 0.1;
    if (vnorm > vlow) {
        if (b_remaining > 0.0) {
            const double rhoe = P.me * P.rho_s + (1.0 - P.me) * P.rho_f;
            const double beta = dmax(1.0 - m, 0.1);
            const double visc = beta * vnorm * 2.0 * P.mu * (1.0 - m) / (dcm_tanh(h + 1.e-2));
            const double gamma = rho * beta * (vnorm * vnorm) * (beta * gz * (coeff * coeff)) /
                                 (dcm_pow(dcm_tanh(h + 1.e-2), 1.0 / 3.0));
            const double t1bot = visc + gamma + tau;
            const double t2top = dmin(t1bot, (1.0 - beta * P.entrainment_rate) * tau);
            const double prat = p / (rho * h);
            double dh = P.entrainment_rate * dt * (t1bot - t2top) / (rhoe * beta * vnorm);
            dh = dmin(dh, b_remaining);
            h = h + dh;
            const double hm = h * m + dh * P.me;
            m = hm / h;
            hchi = hchi + dh * P.chie;
            const double hrho = hm * P.rho_s + (h - hm) * P.rho_f;
            rho = hrho / h;
            b_eroded = b_eroded + dh;
            p = prat * hrho;
        }
    }
}

// src2 body for one interior cell (src2.f90:100-300; bed_normal and curvature
// off).  Updates the six state components and bdif in place; h_e is the
// erodible depth aux(i_ent) (read only with entrainment on).
DC_HD void src2_cell(const dclaw_params& P, double dt, double& qh, double& qhu, double& qhv,
                     double& qhm, double& qp, double& qhchi, double& qbdif, double h_e)
{
    const double drytol = P.dry_tolerance;
    double h = qh;
    if (h <= drytol) {
        qh = 0.0; qhu = 0.0; qhv = 0.0; qhm = 0.0; qp = 0.0; qhchi = 0.0;
        return;
    }
    double hu = qhu, hv = qhv, hm = qhm, p = qp, hchi = qhchi;
    const double rhoh = hm * P.rho_s + (h - hm) * P.rho_f;
    if (hm < 0.0) return;
    const double gz = P.grav, gx = 0.0;
    double u, v, m, chi, rho;
    qfix(P, h, hu, hv, hm, p, hchi, u, v, m, chi, rho, gz);
    Vars sv = setvars(P, h, u, v, m, p, chi, gz);
    const double hvnorm0 = sqrt(hu * hu + hv * hv);
    double vnorm = hvnorm0 / h;
    if (hvnorm0 > 0.0) {
        vnorm = dmax(0.0, vnorm - dt * sv.tau / rhoh);
        vnorm = vnorm * dcm_exp(-((1.0 - m) * 2.0 * P.mu * dt / (h * rhoh)));
        const double hvnorm = h * vnorm;
        hu = hvnorm * hu / hvnorm0 + gx * h * dt;
        hv = hvnorm * hv / hvnorm0;
        u = hu / h;
        v = hv / h;
    }
    // setvars here only recomputes derived quantities (src2.f90:185)
    if (P.src2method == -1) {
        hu = h * u; hv = h * v; hm = h * m; hchi = h * chi;
    } else {
        mp_update_relax(P, dt, h, u, v, m, p, chi, rhoh, gz);
        hu = h * u; hv = h * v; hm = h * m; hchi = h * chi;
    }
    if (h <= drytol) {
        qh = 0.0; qhu = 0.0; qhv = 0.0; qhm = 0.0; qp = 0.0; qhchi = 0.0;
        return;
    }
    if (P.entrainment == 1) {
        // src2.f90:250-263: setvars after the source integration feeds
        // ent_dclaw4; h, m, p, hchi and bdif change there, while hu, hv and
        // hm keep their values from before (they are not recomputed)
        const Vars s3 = setvars(P, h, u, v, m, p, chi, gz);
        rho = s3.rho;
        double b_eroded = qbdif;
        const double b_remaining = h_e - b_eroded;
        if (P.entrainment_method == 0) {
            ent_dclaw4(P, dt, h, u, v, m, p, rho, hchi, gz, s3.tau, b_eroded, b_remaining);
            qbdif = b_eroded;
        }
    }
    if (P.friction_forcing && P.manning_coefficient > 0.0) {
        if (h <= P.friction_depth) {
            double coeffmanning = P.manning_coefficient;
            if (P.dd_manning)
                coeffmanning = P.manning_max + (P.manning_coefficient - P.manning_max) * 0.5
                               * (1.0 + dcm_tanh(300.0 * (h - 0.05)));
            const double beta = 1.0;
            const double gamma = beta * sqrt(hu * hu + hv * hv) * (gz * (coeffmanning * coeffmanning))
                                 / dcm_pow(h, 7.0 / 3.0);
            const double dgamma = 1.0 + dt * gamma;
            hu = hu / dgamma;
            hv = hv / dgamma;
            qfix(P, h, hu, hv, hm, p, hchi, u, v, m, chi, rho, gz);
        }
    }
    qh = h; qhu = hu; qhv = hv; qhm = hm; qp = p; qhchi = hchi;
}

// calc_taudir body for one cell (digclaw_module.f90:546-713, bed_normal = 0).
// c, l, r, bo, t: Q0 states (after b4step2's qfix) of the centre and its
// left/right/bottom/top neighbours; *_b: aux(1) - bdif of each.
struct Taudir {
    double fsphi, tx, ty;
};

DC_HD Taudir calc_taudir_cell(const dclaw_params& P, double dx, double dy, const Cell& C,
                              double bC, const Cell& L, double bL_, const Cell& R, double bR_,
                              const Cell& B, double bB_, const Cell& T, double bT_,
                              const RhoTau& svC)
{
    const double gz = P.grav, drytol = P.dry_tolerance;
    double u, v, m, chi, rho, tu, tv, tm, tchi, trho;
    double h = C.h, hu = C.hu, hv = C.hv, hm = C.hm, p = C.p, hchi = C.hchi;
    qfix(P, h, hu, hv, hm, p, hchi, u, v, m, chi, rho, gz);
    double eta = h + bC;

    double hL = L.h, a1 = L.hu, a2 = L.hv, a3 = L.hm, a4 = L.p, a5 = L.hchi;
    qfix(P, hL, a1, a2, a3, a4, a5, tu, tv, tm, tchi, trho, gz);
    double etaL = hL + bL_;
    if (hL < drytol) etaL = dmin(etaL, eta);

    double hR = R.h; a1 = R.hu; a2 = R.hv; a3 = R.hm; a4 = R.p; a5 = R.hchi;
    qfix(P, hR, a1, a2, a3, a4, a5, tu, tv, tm, tchi, trho, gz);
    double etaR = hR + bR_;
    if (hR < drytol) etaR = dmin(etaR, eta);

    double hB = B.h; a1 = B.hu; a2 = B.hv; a3 = B.hm; a4 = B.p; a5 = B.hchi;
    qfix(P, hB, a1, a2, a3, a4, a5, tu, tv, tm, tchi, trho, gz);
    double etaB = hB + bB_;
    if (hB < drytol) etaB = dmin(etaB, eta);

    double hT = T.h; a1 = T.hu; a2 = T.hv; a3 = T.hm; a4 = T.p; a5 = T.hchi;
    qfix(P, hT, a1, a2, a3, a4, a5, tu, tv, tm, tchi, trho, gz);
    double etaT = hT + bT_;
    if (hT < drytol) etaT = dmin(etaT, eta);

    if (h < drytol) {
        eta = dmin(etaL, eta);
        eta = dmin(etaB, eta);
    }
    Taudir o;
    if ((h + hL + hB + hR + hT) < drytol) {
        o.tx = 0.0; o.ty = 0.0; o.fsphi = 0.0;
        return o;
    }
    // setvars of the centre's qfix'ed state (= F1 of the cell), read only
    // where hvnorm > 0, i.e. on a wet centre
    (void)m;
    (void)chi;
    rho = svC.rho;
    const double sth = sin(0.0);
    const double FxL = -(gz * 0.5 * (h + hL) * (eta - etaL) / dx) + gz * 0.5 * (h + hL) * sth;
    const double FyL = -(gz * 0.5 * (h + hB) * (eta - etaB) / dy);
    const double FxR = -(gz * 0.5 * (h + hR) * (etaR - eta) / dx) + gz * 0.5 * (h + hR) * sth;
    const double FyR = -(gz * 0.5 * (h + hT) * (etaT - eta) / dy);
    double Fx, Fy;
    if (FxL * FxR > 0.0) Fx = dsign(dmin(fabs(FxL), fabs(FxR)), FxL);
    else Fx = 0.0;
    if (FyL * FyR > 0.0) Fy = dsign(dmin(fabs(FyL), fabs(FyR)), FyL);
    else Fy = 0.0;

    const double hvnorm = sqrt(hu * hu + hv * hv);
    if (hvnorm > 0.0) {
        o.tx = -(dx * hu / hvnorm);
        o.ty = -(dy * hv / hvnorm);
        const double dot = dmin(dmax(0.0, Fx * hu), dmax(0.0, Fy * hv));
        if (dot > 0.0) {
            const double Fproj = dot / hvnorm;
            o.fsphi = dmin(1.0, Fproj * rho / dmax(svC.tau, 1e-16));
        } else {
            o.fsphi = 0.0;
        }
    } else {
        if ((FxL * FxL + Fy * Fy) > 0.0) o.tx = -(dx * FxL / sqrt(FxL * FxL + Fy * Fy));
        else o.tx = dx * 1.0;
        if ((Fx * Fx + FyL * FyL) > 0.0) o.ty = -(dy * FyL / sqrt(Fx * Fx + FyL * FyL));
        else o.ty = dy * 1.0;
        if ((o.ty * o.ty + o.tx * o.tx) > 0.0) o.fsphi = 1.0;
        else o.fsphi = 0.0;
    }
    return o;
}

}  // namespace dclaw
