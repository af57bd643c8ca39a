/*
 * dclaw_oracle.c -- CPU restatement of D-Claw's per-patch finite-volume step.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle of the tests).  Never linked into the
 * product.
 *
 * Each routine restates the reference Fortran it cites, statement by
 * statement and in the same floating-point evaluation order (Fortran
 * left-to-right, `x**2` as x*x, unary minus outside the product), compiled
 * with -ffp-contract=off and glibc libm, so that it reproduces the Fortran's
 * arithmetic.  MAX/MIN are written as the x86 maxsd/minsd semantics
 * ((a > b) ? a : b), SIGN(a,b) as copysign(|a|, b).
 *
 * Routines owned by D-Claw follow the files in src/2d/dig.  The GeoClaw/AMRClaw glue
 * that is not part of the reference tree (step2, the inline MC limiter,
 * stepgrid's update, bc2amr extrap/wall, the dt rule) is restated from
 * Clawpack 5's published algorithm; those parts are pinned by invariants
 * (conservation, lake at rest, symmetry, closed forms, observed order of
 * accuracy), not by reference source.  The physics of every routine is
 * pinned by consequences of the governing equations (tests/test_oracle_*.py),
 * and tests/test_oracle_mutants.py checks that those pins reject altered
 * copies of this file.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "dclaw_oracle.h"

/* Elementary functions.  Default: the portable deterministic implementations
 * of dcmath.h (the same source the GPU step uses), so that a CPU run and a
 * GPU run of the scheme can agree bit for bit over thousands of steps.
 * With -DORACLE_GLIBC: glibc libm, as a gfortran build of the reference
 * would call; the tests use that build to bound the effect of the libm. */
#ifdef ORACLE_GLIBC
#define D_EXP exp
#define D_ATAN atan
#define D_TAN tan
#define D_TANH tanh
#define D_POW pow
#define D_LOG log
#else
#define DCM_FN static inline
#include "../dclaw_amd/csrc/dcmath.h"
#define D_EXP dcm_exp
#define D_ATAN dcm_atan
#define D_TAN dcm_tan
#define D_TANH dcm_tanh
#define D_POW dcm_pow
#define D_LOG dcm_log
#endif

#define MEQN 7
#define I_H 1
#define I_HU 2
#define I_HV 3
#define I_HM 4
#define I_PB 5
#define I_HCHI 6
#define I_BDIF 7
#define I_PHI 2
#define I_THETA 3
#define I_FSPHI 4
#define I_TAUDIR_X 5
#define I_TAUDIR_Y 6

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline double dsign(double a, double b) { return copysign(fabs(a), b); }

/* 1-D slice element (m 1-based, i in 1-mbc..) */
#define S1(arr, ncomp, m, i) (arr)[((m) - 1) + (ncomp) * ((i) + mbc - 1)]
/* 2-D patch element */
#define P2(arr, ncomp, m, i, j) \
    (arr)[((m) - 1) + (size_t)(ncomp) * ((size_t)((i) + mbc - 1) + (size_t)((j) + mbc - 1) * (size_t)(mx + 2 * mbc))]

/* ------------------------------------------------------------------------ */
/* qfix: digclaw_module.f90:285-345 */
static void qfix(const oracle_params* P, double* h, double* hu, double* hv, double* hm,
                 double* p, double* hchi, double* u, double* v, double* m, double* chi,
                 double* rho, double gz)
{
    if (*h <= P->dry_tolerance) {
        *h = 0.0; *hu = 0.0; *hv = 0.0; *hm = 0.0; *p = 0.0; *hchi = 0.0;
        *u = 0.0; *v = 0.0; *m = 0.0; *chi = 0.0; *rho = 0.0;
        return;
    }
    *u = *hu / *h;
    *v = *hv / *h;
    *m = *hm / *h;
    *chi = *hchi / *h;
    *m = dmax(*m, 0.0);
    *m = dmin(*m, 1.0);
    *hm = *h * *m;
    *chi = dmax(*chi, 0.0);
    *chi = dmin(*chi, 1.0);
    *hchi = *h * *chi;
    *rho = P->rho_s * *m + (1.0 - *m) * P->rho_f;
    double pmax = *rho * gz * *h;
    *p = dmax(0.0, *p);
    *p = dmin(pmax, *p);
}

/* qfix_cmass: digclaw_module.f90:353-387 */
static void qfix_cmass(const oracle_params* P, double* h, double* m, double* p, double* rho,
                       double* p_exc, double* hu, double* hv, double* hm, double u, double v,
                       double rhoh, double gz)
{
    *m = dmax(*m, 0.0);
    *m = dmin(*m, 1.0);
    *rho = *m * (P->rho_s - P->rho_f) + P->rho_f;
    *h = rhoh / *rho;
    double p_exc_max = rhoh * gz - P->rho_f * *h * gz;
    double p_exc_min = 0.0 - P->rho_f * *h * gz;
    *p_exc = dmax(p_exc_min, *p_exc);
    *p_exc = dmin(p_exc_max, *p_exc);
    *p = P->rho_f * gz * *h + *p_exc;
    *hu = *h * u;
    *hv = *h * v;
    *hm = *h * *m;
}

/* setvars: digclaw_module.f90:394-491.  `phi` is the module global that
 * calc_taudir overwrites (digclaw_module.f90:557); callers pass P with it. */
static void setvars(const oracle_params* P, double h, double u, double v, double m, double p,
                    double chi, double gz, double* rho, double* kperm, double* alphainv,
                    double* m_eq, double* tanpsi, double* tau)
{
    *rho = m * (P->rho_s - P->rho_f) + P->rho_f;
    double kr_chi = 1.0;
    if (P->segregation == 1) kr_chi = D_POW(10.0, P->delta_kr_order * 2.0 * (chi - 0.5));
    *kperm = kr_chi * P->kref * D_EXP(-((m - P->mref) / 0.04));
    double sig_eff = dmax(0.0, *rho * gz * h - p);
    double sig_0 = P->sigma_0;
    if (P->alphamethod == 1)
        sig_0 = 0.50 * P->alpha_c * P->rho_f * (P->rho_s - P->rho_f) * gz * h / *rho;
    *alphainv = m * (sig_eff + sig_0) / P->alpha_c;
    double vnorm = sqrt(u * u + v * v);
    double shear;
    if (h > P->dry_tolerance) shear = 2.0 * vnorm / h;
    else shear = 0.0;
    double sd = shear * P->delta;
    double Nden = P->rho_s * (sd * sd) + sig_eff;
    double Nnum = P->mu * shear;
    if (Nnum <= 0.0) *m_eq = P->m_crit;
    else *m_eq = P->m_crit * (sqrt(Nden) / (sqrt(Nden) + sqrt(Nnum)));
    *tanpsi = P->c1 * (m - *m_eq);
    double psi = D_ATAN(P->c1 * (m - *m_eq));
    *tau = sig_eff * D_TAN(P->phi + psi);
    *tau = *tau * 0.5 * (1.0 + D_TANH(100.0 * (m - 0.1)));
}

/* riemanntype: riemannsolvers_dclaw.f:526-637 */
static void riemanntype(double hL, double hR, double uL, double uR, double* hm, double* s1m,
                        double* s2m, int* rare1, int* rare2, int maxiter, double drytol,
                        double g)
{
    double um, u1m, u2m, delu, h_max, h_min, h0, F_max, F_min, dfdh, F0, slope, gL, gR;
    h_min = dmin(hR, hL);
    h_max = dmax(hR, hL);
    delu = uR - uL;
    if (h_min <= drytol) {
        *hm = 0.0;
        um = 0.0;
        *s1m = uR + uL - 2.0 * sqrt(g * hR) + 2.0 * sqrt(g * hL);
        *s2m = uR + uL - 2.0 * sqrt(g * hR) + 2.0 * sqrt(g * hL);
        if (hL <= 0.0) { *rare2 = 1; *rare1 = 0; }
        else { *rare1 = 1; *rare2 = 0; }
    } else {
        F_min = delu + 2.0 * (sqrt(g * h_min) - sqrt(g * h_max));
        F_max = delu + (h_max - h_min) * (sqrt(.5 * g * (h_max + h_min) / (h_max * h_min)));
        if (F_min > 0.0) {
            double t = dmax(0.0, -delu + 2.0 * (sqrt(g * hL) + sqrt(g * hR)));
            *hm = (1.0 / (16.0 * g)) * (t * t);
            um = dsign(1.0, *hm) * (uL + 2.0 * (sqrt(g * hL) - sqrt(g * *hm)));
            *s1m = uL + 2.0 * sqrt(g * hL) - 3.0 * sqrt(g * *hm);
            *s2m = uR - 2.0 * sqrt(g * hR) + 3.0 * sqrt(g * *hm);
            *rare1 = 1;
            *rare2 = 1;
        } else if (F_max <= 0.0) {
            h0 = h_max;
            for (int iter = 1; iter <= maxiter; iter++) {
                gL = sqrt(.5 * g * (1.0 / h0 + 1.0 / hL));
                gR = sqrt(.5 * g * (1.0 / h0 + 1.0 / hR));
                F0 = delu + (h0 - hL) * gL + (h0 - hR) * gR;
                dfdh = gL - g * (h0 - hL) / (4.0 * (h0 * h0) * gL) + gR
                       - g * (h0 - hR) / (4.0 * (h0 * h0) * gR);
                slope = 2.0 * sqrt(h0) * dfdh;
                double t = sqrt(h0) - F0 / slope;
                h0 = t * t;
            }
            *hm = h0;
            u1m = uL - (*hm - hL) * sqrt((.5 * g) * (1.0 / *hm + 1.0 / hL));
            u2m = uR + (*hm - hR) * sqrt((.5 * g) * (1.0 / *hm + 1.0 / hR));
            um = .5 * (u1m + u2m);
            *s1m = u1m - sqrt(g * *hm);
            *s2m = u2m + sqrt(g * *hm);
            *rare1 = 0;
            *rare2 = 0;
        } else {
            h0 = h_min;
            for (int iter = 1; iter <= maxiter; iter++) {
                F0 = delu + 2.0 * (sqrt(g * h0) - sqrt(g * h_max))
                     + (h0 - h_min) * sqrt(.5 * g * (1.0 / h0 + 1.0 / h_min));
                slope = (F_max - F0) / (h_max - h_min);
                h0 = h0 - F0 / slope;
            }
            *hm = h0;
            if (hL > hR) {
                um = uL + 2.0 * sqrt(g * hL) - 2.0 * sqrt(g * *hm);
                *s1m = uL + 2.0 * sqrt(g * hL) - 3.0 * sqrt(g * *hm);
                *s2m = uL + 2.0 * sqrt(g * hL) - sqrt(g * *hm);
                *rare1 = 1;
                *rare2 = 0;
            } else {
                *s2m = uR - 2.0 * sqrt(g * hR) + 3.0 * sqrt(g * *hm);
                *s1m = uR - 2.0 * sqrt(g * hR) + sqrt(g * *hm);
                um = uR - 2.0 * sqrt(g * hR) + 2.0 * sqrt(g * *hm);
                *rare2 = 1;
                *rare1 = 0;
            }
        }
    }
    (void)um;
}

/* riemann_dig2_aug_sswave_ez: riemannsolvers_dclaw.f:3-521.
 * fw is fw(6,3) column-major: fw[(m-1) + 6*(mw-1)]. */
#define FW(m, mw) fw[((m) - 1) + 6 * ((mw) - 1)]
static void riemann_dig2_aug_sswave_ez(const oracle_params* P, int ixy, double hL, double hR,
        double huL, double huR, double hvL, double hvR, double hmL, double hmR, double pL,
        double pR, double bL, double bR, double uL, double uR, double vL, double vR,
        double mL, double mR, double thetaL, double thetaR, double* sw, double* fw,
        int wallprob, double taudirL, double taudirR, double chiL, double chiR, double fsL,
        double fsR)
{
    (void)ixy;
    const double grav = P->grav, rho_f = P->rho_f, kappa = P->kappa;
    double h, u, v, mbar, hustar_interface, u_interface;
    double R[3][4];   /* R(0:2,1:3) */
    double del[5];    /* del(0:4) */
    double beta[4];   /* beta(1:3) */
    double rho, rhoL, rhoR, tauL, tauR, tau, gzL, gzR, tanpsi, kperm, m_eq, alphainv;
    double theta, gamma, eps, sL, sR, sRoe1, sRoe2, sE1, sE2, uhat, chat;
    double delb, s1m, s2m, hm, criticaltol, criticaltol_2;
    double s1s2bar, s1s2tilde, hbar, source2dx, veltol2, hstarHLL, deldelh, drytol, gz, geps;
    double tausource, gammaL, gammaR, theta1, theta3, vnorm, alpha_seg, a, b, c;
    double seg_L, seg_R;
    int sonic, rare1, rare2, rarecorrector, static_, cwavetype;
    const int delta_huvm_revert = 1;

    veltol2 = 0.0;
    drytol = P->dry_tolerance;
    criticaltol = dmax(drytol * grav, 1e-6);
    criticaltol_2 = sqrt(criticaltol);

    gzL = grav * cos(thetaL);
    gzR = grav * cos(thetaR);
    theta = 0.5 * (thetaL + thetaR);
    gz = grav * cos(theta);
    alpha_seg = 1.0 - P->beta_seg;

    if (hL >= drytol && hR >= drytol) {
        setvars(P, hL, uL, vL, mL, pL, chiL, gzL, &rhoL, &kperm, &alphainv, &m_eq, &tanpsi, &tauL);
        setvars(P, hR, uR, vR, mR, pR, chiR, gzR, &rhoR, &kperm, &alphainv, &m_eq, &tanpsi, &tauR);
        h = 0.5 * (hL + hR);
        v = 0.5 * (vL + vR);
        mbar = 0.5 * (mL + mR);
    } else if (hL >= drytol) {
        setvars(P, hL, uL, vL, mL, pL, chiL, gzL, &rhoL, &kperm, &alphainv, &m_eq, &tanpsi, &tauL);
        setvars(P, hL, uL, vL, mL, pL, chiL, gzL, &rhoR, &kperm, &alphainv, &m_eq, &tanpsi, &tauR);
        tauR = 0.5 * tauL;
        h = 0.5 * hL;
        v = vL;
        mbar = mL;
    } else {
        setvars(P, hR, uR, vR, mR, pR, chiR, gzR, &rhoL, &kperm, &alphainv, &m_eq, &tanpsi, &tauL);
        setvars(P, hR, uR, vR, mR, pR, chiR, gzR, &rhoR, &kperm, &alphainv, &m_eq, &tanpsi, &tauR);
        tauL = 0.5 * tauR;
        h = 0.5 * hR;
        v = vR;
        mbar = mR;
    }
    (void)v; (void)mbar;

    tauL = fsL * tauL;
    tauR = fsR * tauR;
    rho = 0.5 * (rhoL + rhoR);
    tau = 0.5 * (tauL + tauR);
    (void)tau;

    if (P->src2method == -1) {
        gamma = rho_f / rho;
        gammaL = rho_f / rhoL;
        gammaR = rho_f / rhoR;
    } else {
        gamma = 0.25 * (rho_f + 3.0 * rho) / rho;
        gammaL = 0.25 * (rho_f + 3.0 * rhoL) / rhoL;
        gammaR = 0.25 * (rho_f + 3.0 * rhoR) / rhoR;
    }

    eps = kappa + (1.0 - kappa) * gamma;
    geps = gz * eps;

    sL = uL - sqrt(geps * hL);
    sR = uR + sqrt(geps * hR);
    uhat = (sqrt(hL) * uL + sqrt(hR) * uR) / (sqrt(hR) + sqrt(hL));
    chat = sqrt(geps * 0.5 * (hR + hL));
    sRoe1 = uhat - chat;
    sRoe2 = uhat + chat;

    sE1 = dmin(sL, sRoe1);
    sE2 = dmax(sR, sRoe2);
    sw[0] = sE1;
    sw[2] = sE2;
    u = uhat;

    riemanntype(hL, hR, uL, uR, &hm, &s1m, &s2m, &rare1, &rare2, 1, drytol, geps);
    sw[0] = dmin(sw[0], s2m);
    sw[2] = dmax(sw[2], s1m);
    sw[1] = 0.5 * (sw[2] + sw[0]);

    hstarHLL = dmax((huL - huR + sE2 * hR - sE1 * hL) / (sE2 - sE1), 0.0);

    rarecorrector = 0;
    if (P->rarecorrectortest) {
        double sdelta = sw[2] - sw[0];
        double raremin = 0.5, raremax = 0.9;
        if (rare1 && sE1 * s1m < 0.0) raremin = 0.2;
        if (rare2 && sE2 * s2m < 0.0) raremin = 0.2;
        if (rare1 || rare2) {
            double rare1st = 3.0 * (sqrt(geps * hL) - sqrt(geps * hm));
            double rare2st = 3.0 * (sqrt(geps * hR) - sqrt(geps * hm));
            if (dmax(rare1st, rare2st) > raremin * sdelta &&
                dmax(rare1st, rare2st) < raremax * sdelta) {
                rarecorrector = 1;
                if (rare1st > rare2st) sw[1] = s1m;
                else if (rare2st > rare1st) sw[1] = s2m;
                else sw[1] = 0.5 * (s1m + s2m);
            }
        }
        if (hstarHLL < dmin(hL, hR) / 5.0) rarecorrector = 0;
    }

    delb = (bR - bL);

    hbar = 0.5 * (hL + hR);
    {
        double t = uL + uR;
        s1s2bar = 0.25 * (t * t) - gz * hbar;
    }
    s1s2tilde = dmax(0.0, uL * uR) - gz * hbar;

    sonic = 0;
    if (fabs(s1s2bar) <= criticaltol) sonic = 1;
    else if (s1s2bar * s1s2tilde <= criticaltol * criticaltol) sonic = 1;
    else if (s1s2bar * sE1 * sE2 <= criticaltol * criticaltol) sonic = 1;
    else if (dmin(fabs(sE1), fabs(sE2)) < criticaltol_2) sonic = 1;
    else if (sE1 < criticaltol_2 && s1m > -criticaltol_2) sonic = 1;
    else if (sE2 > -criticaltol_2 && s2m < criticaltol_2) sonic = 1;
    else if ((uL + sqrt(geps * hL)) * (uR + sqrt(geps * hR)) < 0.0) sonic = 1;
    else if ((uL - sqrt(geps * hL)) * (uR - sqrt(geps * hR)) < 0.0) sonic = 1;

    if (sonic) source2dx = -(gz * hbar * delb);
    else source2dx = -(gz * hbar * delb * s1s2tilde / s1s2bar);

    source2dx = dmin(source2dx, gz * dmax(-(hL * delb), -(hR * delb)));
    source2dx = dmax(source2dx, gz * dmin(-(hL * delb), -(hR * delb)));

    if (fabs(u) <= veltol2) source2dx = -(hbar * gz * delb);

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

    R[0][2] = 0.0; R[1][2] = 0.0; R[2][2] = 1.0;
    cwavetype = 1;
    R[0][1] = 1.0; R[1][1] = sw[0]; R[2][1] = sw[0] * sw[0];
    R[0][3] = 1.0; R[1][3] = sw[2]; R[2][3] = sw[2] * sw[2];
    if (rarecorrector) {
        R[0][2] = 1.0; R[1][2] = sw[1]; R[2][2] = sw[1] * sw[1];
        cwavetype = 2;
    }

    del[0] = hR - hL - deldelh;
    del[1] = huR - huL;
    del[2] = hR * (uR * uR) + 0.5 * kappa * gz * (hR * hR)
             - (hL * (uL * uL) + 0.5 * kappa * gz * (hL * hL));
    del[2] = del[2] + (1.0 - kappa) * h * (pR - pL) / rho;
    del[3] = pR - pL - gamma * rho * gz * deldelh;
    del[4] = -(gamma * rho * gz * u * (hR - hL)) + gamma * rho * gz * del[1] + u * (pR - pL);

    vnorm = sqrt(uR * uR + uL * uL + vR * vR + vL * vL);
    static_ = 0;
    if (vnorm > 0.0) {
        tausource = 0.0;
    } else if (0.5 * fabs(taudirR * tauR / rhoR + tauL * taudirR / rhoL) > fabs(del[2] - source2dx)) {
        tausource = del[2] - source2dx;
        del[1] = 0.0;
        del[0] = 0.0;
        del[3] = 0.0;
        del[4] = 0.0;
        static_ = 1;
    } else {
        tausource = 0.5 * ((taudirR * tauR / rhoR) + (tauL * taudirR / rhoL));
        tausource = dsign(tausource, del[2] - source2dx);
    }
    if (wallprob) tausource = 0.0;

    del[2] = del[2] - source2dx - tausource;

    if (sw[0] >= 0.0) { theta1 = thetaR; theta3 = thetaR; }
    else if (sw[2] <= 0.0) { theta1 = thetaL; theta3 = thetaL; }
    else if (sw[1] >= 0.0) { theta1 = thetaL; theta3 = thetaR; }
    else { theta1 = thetaL; theta3 = thetaR; }

    a = sw[0];
    b = sw[1];
    c = sw[2];
    if (cwavetype == 1) {
        beta[1] = (c * del[0] - del[1]) / (c - a);
        beta[2] = a * c * del[0] - (a + c) * del[1] + del[2];
        beta[3] = (del[1] - a * del[0]) / (c - a);
    } else {
        beta[1] = (b * c * del[0] - (b + c) * del[1] + del[2]) / (a * a - a * b - a * c + b * c);
        beta[2] = (-(a * c * del[0]) + (a + c) * del[1] - del[2]) / (a * b - a * c - b * b + b * c);
        beta[3] = (a * b * del[0] - (a + b) * del[1] + del[2]) / (a * b - a * c - b * c + c * c);
    }

    for (int mw = 1; mw <= 3; mw++)
        for (int m = 1; m <= 2; m++) FW(m, mw) = beta[mw] * R[m][mw];

    FW(4, 1) = FW(1, 1) * mL;
    FW(4, 3) = FW(1, 3) * mR;
    FW(4, 2) = hmR * uR - hmL * uL - FW(4, 1) - FW(4, 3);

    FW(3, 1) = FW(1, 1) * vL;
    FW(3, 3) = FW(1, 3) * vR;
    FW(3, 2) = hvR * uR - hvL * uL - FW(3, 1) - FW(3, 3);

    FW(5, 1) = FW(1, 1) * gammaL * rhoL * grav * cos(theta1);
    FW(5, 3) = FW(1, 3) * gammaR * rhoR * grav * cos(theta3);
    FW(5, 2) = del[4] - FW(5, 3) - FW(5, 1);

    seg_L = chiL * hL * uL * (1.0 + (1.0 - alpha_seg) * (1.0 - chiL));
    seg_R = chiR * hR * uR * (1.0 + (1.0 - alpha_seg) * (1.0 - chiR));
    FW(6, 1) = FW(1, 1) * chiL * (1.0 + (1.0 - alpha_seg) * (1.0 - chiL));
    FW(6, 3) = FW(1, 3) * chiR * (1.0 + (1.0 - alpha_seg) * (1.0 - chiR));
    FW(6, 2) = seg_R - seg_L - FW(6, 1) - FW(6, 3);

    if (delta_huvm_revert) {
        hustar_interface = huL + FW(1, 1);
        if (!static_) {
            u_interface = hustar_interface / hstarHLL;
            del[4] = -(gamma * rho * gz * u_interface * (hR - hL)) + gamma * rho * gz * del[1]
                     + u_interface * (pR - pL);
            FW(5, 2) = del[4] - FW(5, 3) - FW(5, 1);
        }
        if (hustar_interface <= 0.0) {
            FW(3, 1) = FW(3, 1) + FW(3, 2);
            FW(4, 1) = FW(4, 1) + FW(4, 2);
            FW(5, 1) = FW(5, 1) + FW(5, 2);
            FW(6, 1) = FW(6, 1) + FW(6, 2);
        } else {
            FW(3, 3) = FW(3, 3) + FW(3, 2);
            FW(4, 3) = FW(4, 3) + FW(4, 2);
            FW(5, 3) = FW(5, 3) + FW(5, 2);
            FW(6, 3) = FW(6, 3) + FW(6, 2);
        }
        FW(3, 2) = 0.0;
        FW(4, 2) = 0.0;
        FW(5, 2) = 0.0;
        FW(6, 2) = 0.0;
    }
}
#undef FW

/* ------------------------------------------------------------------------ */
/* rpn2: rpn2_dclaw.f:24-392 (mcapa = 0, bed_normal = 0).  ql = qr = q1d and
 * qfix is applied to q1d IN PLACE (rpn2_dclaw.f:164-168), exactly as the
 * reference does through its by-reference arguments. */
void oracle_rpn2(const oracle_params* P, int ixy, int mx, int mbc, int maux, double* q1d,
                 const double* aux2, double* fwave, double* s, double* amdq, double* apdq)
{
    const int mwaves = P->mwaves;
    const double drytol = P->dry_tolerance, gmod = P->grav;
    int mhu, nhv;
    if (ixy == 1) { mhu = I_HU; nhv = I_HV; } else { mhu = I_HV; nhv = I_HU; }
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        double wall[3], fw[18], sw[3];
        for (int mw = 1; mw <= mwaves; mw++) {
            s[(mw - 1) + mwaves * (i + mbc - 1)] = 0.0;
            for (int m = 1; m <= MEQN; m++) fwave[(m - 1) + MEQN * ((mw - 1) + mwaves * (i + mbc - 1))] = 0.0;
        }
        for (int k = 0; k < 3; k++) sw[k] = 0.0;
        for (int k = 0; k < 18; k++) fw[k] = 0.0;

        if (S1(q1d, MEQN, I_H, i - 1) <= drytol && S1(q1d, MEQN, I_H, i) <= drytol) continue;

        double taudirR, taudirL;
        if (ixy == 1) {
            taudirR = S1(aux2, maux, I_TAUDIR_X, i);
            taudirL = S1(aux2, maux, I_TAUDIR_X, i - 1);
        } else {
            taudirR = S1(aux2, maux, I_TAUDIR_Y, i);
            taudirL = S1(aux2, maux, I_TAUDIR_Y, i - 1);
        }
        double fsL = S1(aux2, maux, I_FSPHI, i - 1);
        double fsR = S1(aux2, maux, I_FSPHI, i);
        double thetaL = 0.0, thetaR = 0.0;
        double gzL = P->grav, gzR = P->grav;

        double uR, vR, mR, chiR, rhoR, uL, vL, mL, chiL, rhoL;
        qfix(P, &S1(q1d, MEQN, I_H, i), &S1(q1d, MEQN, mhu, i), &S1(q1d, MEQN, nhv, i),
             &S1(q1d, MEQN, I_HM, i), &S1(q1d, MEQN, I_PB, i), &S1(q1d, MEQN, I_HCHI, i),
             &uR, &vR, &mR, &chiR, &rhoR, gzR);
        qfix(P, &S1(q1d, MEQN, I_H, i - 1), &S1(q1d, MEQN, mhu, i - 1), &S1(q1d, MEQN, nhv, i - 1),
             &S1(q1d, MEQN, I_HM, i - 1), &S1(q1d, MEQN, I_PB, i - 1), &S1(q1d, MEQN, I_HCHI, i - 1),
             &uL, &vL, &mL, &chiL, &rhoL, gzL);

        double hL = S1(q1d, MEQN, I_H, i - 1), hR = S1(q1d, MEQN, I_H, i);
        double huL = S1(q1d, MEQN, mhu, i - 1), huR = S1(q1d, MEQN, mhu, i);
        double hvL = S1(q1d, MEQN, nhv, i - 1), hvR = S1(q1d, MEQN, nhv, i);
        double hmL = S1(q1d, MEQN, I_HM, i - 1), hmR = S1(q1d, MEQN, I_HM, i);
        double pL = S1(q1d, MEQN, I_PB, i - 1), pR = S1(q1d, MEQN, I_PB, i);
        double bL = S1(aux2, maux, 1, i - 1) - S1(q1d, MEQN, I_BDIF, i - 1);
        double bR = S1(aux2, maux, 1, i) - S1(q1d, MEQN, I_BDIF, i);
        double chiHL = S1(q1d, MEQN, I_HCHI, i - 1), chiHR = S1(q1d, MEQN, I_HCHI, i);
        (void)chiHL; (void)chiHR;

        for (int mw = 0; mw < 3; mw++) wall[mw] = 1.0;
        int wallprob = 0;
        double hstar, hstartest, s1m, s2m;
        int rare1, rare2;
        if (hR <= drytol) {
            hR = 0.0;
            pR = 0.0;
            mR = mL;
            chiR = chiL;
            riemanntype(hL, hL, uL, -uL, &hstar, &s1m, &s2m, &rare1, &rare2, 1, drytol, gmod);
            hstartest = dmax(hL, hstar);
            if (hstartest + bL < bR) {
                wall[1] = 0.0;
                wall[2] = 0.0;
                wallprob = 1;
                hR = hL; huR = -huL; hvR = hvL; hmR = hmL; bR = bL;
                uR = -uL; vR = vL; mR = mL; pR = pL; chiHR = chiHL; chiR = chiL;
            }
        } else if (hL <= drytol) {
            hL = 0.0;
            pL = 0.0;
            mL = mR;
            chiL = chiR;
            riemanntype(hR, hR, -uR, uR, &hstar, &s1m, &s2m, &rare1, &rare2, 1, drytol, gmod);
            hstartest = dmax(hR, hstar);
            if (hstartest + bR < bL) {
                wall[0] = 0.0;
                wall[1] = 0.0;
                wallprob = 1;
                hL = hR; huL = -huR; hvL = hvR; hmL = hmR; mL = mR; bL = bR;
                uL = -uR; vL = vR; pL = pR; chiHL = chiHR; chiL = chiR;
            }
        }

        riemann_dig2_aug_sswave_ez(P, ixy, hL, hR, huL, huR, hvL, hvR, hmL, hmR, pL, pR, bL,
                                   bR, uL, uR, vL, vR, mL, mR, thetaL, thetaR, sw, fw,
                                   wallprob, taudirL, taudirR, chiL, chiR, fsL, fsR);

        for (int mw = 0; mw < 3; mw++) {
            sw[mw] = sw[mw] * wall[mw];
            for (int m = 0; m < 6; m++) fw[m + 6 * mw] = fw[m + 6 * mw] * wall[mw];
        }

#define SS(mw) s[((mw) - 1) + mwaves * (i + mbc - 1)]
#define FWV(m, mw) fwave[((m) - 1) + MEQN * (((mw) - 1) + mwaves * (i + mbc - 1))]
#define F6(m, mw) fw[((m) - 1) + 6 * ((mw) - 1)]
        SS(1) = sw[0];
        SS(2) = sw[1];
        SS(3) = sw[1];
        SS(4) = sw[1];
        SS(5) = sw[2];

        FWV(I_H, 1) = F6(1, 1);
        FWV(mhu, 1) = F6(2, 1);
        FWV(nhv, 1) = F6(3, 1);
        FWV(I_HM, 1) = F6(4, 1);
        FWV(I_PB, 1) = F6(5, 1);
        FWV(I_HCHI, 1) = F6(6, 1);

        FWV(I_H, 5) = F6(1, 3);
        FWV(mhu, 5) = F6(2, 3);
        FWV(nhv, 5) = F6(3, 3);
        FWV(I_HM, 5) = F6(4, 3);
        FWV(I_PB, 5) = F6(5, 3);
        FWV(I_HCHI, 5) = F6(6, 3);

        FWV(I_H, 2) = F6(1, 2);
        FWV(mhu, 2) = F6(2, 2);
        FWV(nhv, 2) = F6(3, 2);
        FWV(I_HM, 2) = 0.0;
        FWV(I_PB, 2) = 0.0;
        FWV(I_HCHI, 2) = F6(6, 2);

        FWV(I_H, 3) = 0.0;
        FWV(mhu, 3) = 0.0;
        FWV(nhv, 3) = 0.0;
        FWV(I_HM, 3) = F6(4, 2);
        FWV(I_PB, 3) = 0.0;
        FWV(I_HCHI, 3) = 0.0;

        FWV(I_H, 4) = 0.0;
        FWV(mhu, 4) = 0.0;
        FWV(nhv, 4) = 0.0;
        FWV(I_HM, 4) = 0.0;
        FWV(I_PB, 4) = F6(5, 2);
        FWV(I_HCHI, 4) = 0.0;
#undef SS
#undef FWV
#undef F6
    }

    /* fluctuations: rpn2_dclaw.f:347-379 (s == 0 contributes to neither) */
    for (int i = 1 - mbc; i <= mx + mbc; i++)
        for (int m = 1; m <= MEQN; m++) {
            S1(amdq, MEQN, m, i) = 0.0;
            S1(apdq, MEQN, m, i) = 0.0;
        }
    for (int i = 2 - mbc; i <= mx + mbc; i++)
        for (int mw = 1; mw <= mwaves; mw++) {
            double sv = s[(mw - 1) + mwaves * (i + mbc - 1)];
            const double* f = &fwave[MEQN * ((mw - 1) + mwaves * (i + mbc - 1))];
            if (sv < 0.0) {
                for (int m = 1; m <= MEQN; m++) S1(amdq, MEQN, m, i) = S1(amdq, MEQN, m, i) + f[m - 1];
            } else if (sv > 0.0) {
                for (int m = 1; m <= MEQN; m++) S1(apdq, MEQN, m, i) = S1(apdq, MEQN, m, i) + f[m - 1];
            }
        }
}

/* rpt2: rpt2_dclaw.f:16-266 (coordinate_system = 1, bed_normal = 0) */
void oracle_rpt2(const oracle_params* P, int ixy, int imp, int mx, int mbc, int maux,
                 const double* q1d, const double* aux1, const double* aux2, const double* aux3,
                 const double* asdq, double* bmasdq, double* bpasdq)
{
    int mu, mv;
    if (ixy == 1) { mu = I_HU; mv = I_HV; } else { mu = I_HV; mv = I_HU; }
    for (int i = 1 - mbc; i <= mx + mbc; i++)
        for (int m = 1; m <= MEQN; m++) {
            S1(bmasdq, MEQN, m, i) = 0.0;
            S1(bpasdq, MEQN, m, i) = 0.0;
        }
    const double g = P->grav;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        double h, u, v, m, p, chi, eta, topo1, topo3;
        int ic = (imp == 1) ? i - 1 : i;
        h = S1(q1d, MEQN, I_H, ic);
        if (h <= P->dry_tolerance) continue;
        u = S1(q1d, MEQN, mu, ic) / h;
        v = S1(q1d, MEQN, mv, ic) / h;
        m = S1(q1d, MEQN, I_HM, ic) / h;
        p = S1(q1d, MEQN, I_PB, ic);
        chi = S1(q1d, MEQN, I_HCHI, ic) / h;
        eta = h + S1(aux2, maux, 1, ic) - S1(q1d, MEQN, I_BDIF, ic);
        topo1 = S1(aux1, maux, 1, ic);
        topo3 = S1(aux3, maux, 1, ic);
        (void)p;

        double rho = m * P->rho_s + (1.0 - m) * P->rho_f;
        double gamma;
        if (P->src2method == -1) gamma = P->rho_f / rho;
        else gamma = 0.25 * (P->rho_f + 3.0 * rho) / rho;

        if (eta < dmin(topo1, topo3)) continue;

        const double dxdcp = 1.0, dxdcm = 1.0;
        double s[4], beta[4], r[7][4];
        s[1] = v - sqrt(g * h);
        s[2] = v;
        s[3] = v + sqrt(g * h);

        double delf1 = S1(asdq, MEQN, I_H, i);
        double delf2 = S1(asdq, MEQN, mu, i);
        double delf3 = S1(asdq, MEQN, mv, i);

        beta[1] = (s[3] * delf1 - delf3) / (s[3] - s[1]);
        beta[2] = -(u * delf1) + delf2;
        beta[3] = (delf3 - s[1] * delf1) / (s[3] - s[1]);

        r[1][1] = 1.0; r[2][1] = u; r[3][1] = s[1]; r[4][1] = m; r[5][1] = gamma * rho * g; r[6][1] = chi;
        r[1][3] = 1.0; r[2][3] = u; r[3][3] = s[3]; r[4][3] = m; r[5][3] = gamma * rho * g; r[6][3] = chi;
        r[1][2] = 0.0; r[2][2] = 1.0; r[3][2] = 0.0; r[4][2] = 0.0; r[5][2] = 0.0; r[6][2] = 0.0;

        const int comp[7] = {0, I_H, mu, mv, I_HM, I_PB, I_HCHI};
        for (int mw = 1; mw <= 3; mw++) {
            if (s[mw] < 0.0 && eta >= topo1) {
                for (int k = 1; k <= 6; k++)
                    S1(bmasdq, MEQN, comp[k], i) = S1(bmasdq, MEQN, comp[k], i) + dxdcm * s[mw] * beta[mw] * r[k][mw];
            } else if (s[mw] > 0.0 && eta >= topo3) {
                for (int k = 1; k <= 6; k++)
                    S1(bpasdq, MEQN, comp[k], i) = S1(bpasdq, MEQN, comp[k], i) + dxdcp * s[mw] * beta[mw] * r[k][mw];
            }
        }
    }
}

/* limiter: AMRClaw inlinelimiter.f (Clawpack 5; not in the reference tree,
 * called at flux2fw.f:188).  Waves are compared with their UNLIMITED
 * neighbours (dotr is formed before wave i is scaled). */
void oracle_limiter(const oracle_params* P, int mx, int mbc, double* fwave, const double* s)
{
    const int mwaves = P->mwaves;
#define WV(m, mw, i) fwave[((m) - 1) + MEQN * (((mw) - 1) + mwaves * ((i) + mbc - 1))]
    for (int mw = 1; mw <= mwaves; mw++) {
        if (P->mthlim[mw - 1] == 0) continue;
        double dotr = 0.0;
        for (int i = 0; i <= mx + 1; i++) {
            double wnorm2 = 0.0;
            double dotl = dotr;
            dotr = 0.0;
            for (int m = 1; m <= MEQN; m++) {
                wnorm2 = wnorm2 + WV(m, mw, i) * WV(m, mw, i);
                dotr = dotr + WV(m, mw, i) * WV(m, mw, i + 1);
            }
            if (i == 0) continue;
            if (wnorm2 == 0.0) continue;
            double r;
            if (s[(mw - 1) + mwaves * (i + mbc - 1)] > 0.0) r = dotl / wnorm2;
            else r = dotr / wnorm2;
            double wlimitr;
            switch (P->mthlim[mw - 1]) {
            case 1: wlimitr = dmax(0.0, dmin(1.0, r)); break;
            case 2: wlimitr = dmax(dmax(0.0, dmin(1.0, 2.0 * r)), dmin(2.0, r)); break;
            case 3: wlimitr = (r + fabs(r)) / (1.0 + fabs(r)); break;
            case 4: {
                double c = (1.0 + r) / 2.0;
                wlimitr = dmax(0.0, dmin(dmin(c, 2.0), 2.0 * r));
                break;
            }
            default: wlimitr = r; break;
            }
            for (int m = 1; m <= MEQN; m++) WV(m, mw, i) = wlimitr * WV(m, mw, i);
        }
    }
#undef WV
}

/* flux2: flux2fw.f:8-286 (relimit = .false., coordinate_system = 1) */
void oracle_flux2(const oracle_params* P, int ixy, int maxm, int mx, int mbc, int maux,
                  double* q1d, double* dtdx1d, const double* aux1, const double* aux2,
                  const double* aux3, double* faddm, double* faddp, double* gaddm,
                  double* gaddp, double* cfl1d)
{
    const int mwaves = P->mwaves;
    const int n = maxm + 2 * mbc;
    double* fwave = (double*)malloc(sizeof(double) * MEQN * mwaves * n);
    double* s = (double*)malloc(sizeof(double) * mwaves * n);
    double* amdq = (double*)malloc(sizeof(double) * MEQN * n);
    double* apdq = (double*)malloc(sizeof(double) * MEQN * n);
    double* cqxx = (double*)calloc((size_t)MEQN * n, sizeof(double));
    double* bmasdq = (double*)malloc(sizeof(double) * MEQN * n);
    double* bpasdq = (double*)malloc(sizeof(double) * MEQN * n);
#define GADD(arr, m, i, side) (arr)[((m) - 1) + MEQN * (((i) + mbc - 1) + n * ((side) - 1))]
    int limit = 0;
    for (int mw = 0; mw < mwaves; mw++) if (P->mthlim[mw] > 0) limit = 1;

    for (int side = 1; side <= 2; side++)
        for (int i = 1 - mbc; i <= mx + mbc; i++)
            for (int m = 1; m <= MEQN; m++) {
                S1(faddm, MEQN, m, i) = 0.0;
                S1(faddp, MEQN, m, i) = 0.0;
                GADD(gaddm, m, i, side) = 0.0;
                GADD(gaddp, m, i, side) = 0.0;
            }

    oracle_rpn2(P, ixy, mx, mbc, maux, q1d, aux2, fwave, s, amdq, apdq);

    for (int i = 1 - mbc + 1; i <= mx + mbc - 1; i++)
        for (int m = 1; m <= MEQN; m++) {
            S1(faddp, MEQN, m, i) = S1(faddp, MEQN, m, i) - S1(apdq, MEQN, m, i);
            S1(faddm, MEQN, m, i) = S1(faddm, MEQN, m, i) + S1(amdq, MEQN, m, i);
        }

    *cfl1d = 0.0;
    for (int mw = 1; mw <= mwaves; mw++)
        for (int i = 1; i <= mx + 1; i++) {
            double sv = s[(mw - 1) + mwaves * (i + mbc - 1)];
            double a = dtdx1d[i + mbc - 1] * sv;
            double b = -(dtdx1d[i - 1 + mbc - 1] * sv);
            /* dmax1(cfl1d, a, b) evaluated left to right */
            *cfl1d = dmax(dmax(*cfl1d, a), b);
        }

    if (P->order != 1) {
        if (limit) oracle_limiter(P, mx, mbc, fwave, s);
        for (int i = 1; i <= mx + 1; i++) {
            dtdx1d[i - 1 + mbc - 1] = 0.5 * (dtdx1d[i - 1 + mbc - 1] + dtdx1d[i + mbc - 1]);
            double hleft2 = S1(q1d, MEQN, I_H, i - 2);
            double hleft = S1(q1d, MEQN, I_H, i - 1);
            double hright = S1(q1d, MEQN, I_H, i);
            double hright2 = S1(q1d, MEQN, I_H, i + 1);
            for (int m = 1; m <= MEQN; m++) {
                S1(cqxx, MEQN, m, i) = 0.0;
                if (hleft * hright * hleft2 * hright2 > 0.0) {
                    for (int mw = 1; mw <= mwaves; mw++) {
                        double sv = s[(mw - 1) + mwaves * (i + mbc - 1)];
                        double fv = fwave[(m - 1) + MEQN * ((mw - 1) + mwaves * (i + mbc - 1))];
                        S1(cqxx, MEQN, m, i) = S1(cqxx, MEQN, m, i)
                            + dsign(1.0, sv) * (1.0 - fabs(sv) * dtdx1d[i - 1 + mbc - 1]) * fv;
                    }
                }
                S1(faddm, MEQN, m, i) = S1(faddm, MEQN, m, i) + 0.5 * S1(cqxx, MEQN, m, i);
                S1(faddp, MEQN, m, i) = S1(faddp, MEQN, m, i) + 0.5 * S1(cqxx, MEQN, m, i);
            }
        }
    }

    if (P->transverse != 0) {
        if (P->order > 1 && P->transverse == 2) {
            for (int i = 1; i <= mx + 1; i++)
                for (int m = 1; m <= MEQN; m++) {
                    S1(amdq, MEQN, m, i) = S1(amdq, MEQN, m, i) + S1(cqxx, MEQN, m, i);
                    S1(apdq, MEQN, m, i) = S1(apdq, MEQN, m, i) - S1(cqxx, MEQN, m, i);
                }
        }
        oracle_rpt2(P, ixy, 1, mx, mbc, maux, q1d, aux1, aux2, aux3, amdq, bmasdq, bpasdq);
        for (int i = 1; i <= mx + 1; i++)
            for (int m = 1; m <= MEQN; m++) {
                double gupdate = 0.5 * dtdx1d[i - 1 + mbc - 1] * S1(bmasdq, MEQN, m, i);
                GADD(gaddm, m, i - 1, 1) = GADD(gaddm, m, i - 1, 1) - gupdate;
                GADD(gaddp, m, i - 1, 1) = GADD(gaddp, m, i - 1, 1) - gupdate;
                gupdate = 0.5 * dtdx1d[i - 1 + mbc - 1] * S1(bpasdq, MEQN, m, i);
                GADD(gaddm, m, i - 1, 2) = GADD(gaddm, m, i - 1, 2) - gupdate;
                GADD(gaddp, m, i - 1, 2) = GADD(gaddp, m, i - 1, 2) - gupdate;
            }
        oracle_rpt2(P, ixy, 2, mx, mbc, maux, q1d, aux1, aux2, aux3, apdq, bmasdq, bpasdq);
        for (int i = 1; i <= mx + 1; i++)
            for (int m = 1; m <= MEQN; m++) {
                double gupdate = 0.5 * dtdx1d[i - 1 + mbc - 1] * S1(bmasdq, MEQN, m, i);
                GADD(gaddm, m, i, 1) = GADD(gaddm, m, i, 1) - gupdate;
                GADD(gaddp, m, i, 1) = GADD(gaddp, m, i, 1) - gupdate;
                gupdate = 0.5 * dtdx1d[i - 1 + mbc - 1] * S1(bpasdq, MEQN, m, i);
                GADD(gaddm, m, i, 2) = GADD(gaddm, m, i, 2) - gupdate;
                GADD(gaddp, m, i, 2) = GADD(gaddp, m, i, 2) - gupdate;
            }
    }
#undef GADD
    free(fwave); free(s); free(amdq); free(apdq); free(cqxx); free(bmasdq); free(bpasdq);
}

/* ------------------------------------------------------------------------ */
/* calc_taudir: digclaw_module.f90:500-719 (bed_normal = 0).  Overwrites the
 * global phi with aux(i_phi,i,j) for every cell it visits (:557). */
void oracle_calc_taudir(oracle_params* P, int mx, int my, int mbc, int maux, double dx,
                        double dy, const double* q, double* aux)
{
    const double gz = P->grav, theta = 0.0, drytol = P->dry_tolerance;
    for (int j = 2 - mbc; j <= my + mbc - 1; j++)
        for (int i = 2 - mbc; i <= mx + mbc - 1; i++) {
            double u, v, m, chi, rho, uL, vL, mL, chiL, rhoL, uR, vR, mR, chiR, rhoR;
            double uB, vB, mB, chiB, rhoB, uT, vT, mT, chiT, rhoT;
            double h = P2(q, MEQN, I_H, i, j), hu = P2(q, MEQN, I_HU, i, j), hv = P2(q, MEQN, I_HV, i, j);
            double hm = P2(q, MEQN, I_HM, i, j), p = P2(q, MEQN, I_PB, i, j), hchi = P2(q, MEQN, I_HCHI, i, j);
            qfix(P, &h, &hu, &hv, &hm, &p, &hchi, &u, &v, &m, &chi, &rho, gz);
            double b = P2(aux, maux, 1, i, j) - P2(q, MEQN, I_BDIF, i, j);
            double eta = h + b;
            P->phi = P2(aux, maux, I_PHI, i, j);

#define NB(SUF, ii, jj)                                                                        \
    double h##SUF = P2(q, MEQN, I_H, ii, jj), hu##SUF = P2(q, MEQN, I_HU, ii, jj);            \
    double hv##SUF = P2(q, MEQN, I_HV, ii, jj), hm##SUF = P2(q, MEQN, I_HM, ii, jj);          \
    double p##SUF = P2(q, MEQN, I_PB, ii, jj), hchi##SUF = P2(q, MEQN, I_HCHI, ii, jj);       \
    qfix(P, &h##SUF, &hu##SUF, &hv##SUF, &hm##SUF, &p##SUF, &hchi##SUF, &u##SUF, &v##SUF,      \
         &m##SUF, &chi##SUF, &rho##SUF, gz);                                                   \
    double b##SUF = P2(aux, maux, 1, ii, jj) - P2(q, MEQN, I_BDIF, ii, jj);                   \
    double eta##SUF = h##SUF + b##SUF;                                                         \
    if (h##SUF < drytol) eta##SUF = dmin(eta##SUF, eta);

            NB(L, i - 1, j)
            NB(R, i + 1, j)
            NB(B, i, j - 1)
            NB(T, i, j + 1)
#undef NB
            (void)huL; (void)hvL; (void)huR; (void)hvR; (void)huB; (void)hvB; (void)huT; (void)hvT;
            (void)bL; (void)bR; (void)bB; (void)bT;

            if (h < drytol) {
                eta = dmin(etaL, eta);
                eta = dmin(etaB, eta);
            }
            if ((h + hL + hB + hR + hT) < drytol) {
                P2(aux, maux, I_TAUDIR_X, i, j) = 0.0;
                P2(aux, maux, I_TAUDIR_Y, i, j) = 0.0;
                P2(aux, maux, I_FSPHI, i, j) = 0.0;
                continue;
            }
            double kperm, alphainv, m_eq, tanpsi, tau;
            setvars(P, h, u, v, m, p, chi, gz, &rho, &kperm, &alphainv, &m_eq, &tanpsi, &tau);

            const double sth = sin(theta);
            double FxL = -(gz * 0.5 * (h + hL) * (eta - etaL) / dx) + gz * 0.5 * (h + hL) * sth;
            double FyL = -(gz * 0.5 * (h + hB) * (eta - etaB) / dy);
            double FxR = -(gz * 0.5 * (h + hR) * (etaR - eta) / dx) + gz * 0.5 * (h + hR) * sth;
            double FyR = -(gz * 0.5 * (h + hT) * (etaT - eta) / dy);
            double Fx, Fy;
            if (FxL * FxR > 0.0) Fx = dsign(dmin(fabs(FxL), fabs(FxR)), FxL);
            else Fx = 0.0;
            if (FyL * FyR > 0.0) Fy = dsign(dmin(fabs(FyL), fabs(FyR)), FyL);
            else Fy = 0.0;

            double hvnorm = sqrt(hu * hu + hv * hv);
            if (hvnorm > 0.0) {
                P2(aux, maux, I_TAUDIR_X, i, j) = -(dx * hu / hvnorm);
                P2(aux, maux, I_TAUDIR_Y, i, j) = -(dy * hv / hvnorm);
                double dot = dmin(dmax(0.0, Fx * hu), dmax(0.0, Fy * hv));
                if (dot > 0.0) {
                    double Fproj = dot / hvnorm;
                    P2(aux, maux, I_FSPHI, i, j) = dmin(1.0, Fproj * rho / dmax(tau, 1e-16));
                } else {
                    P2(aux, maux, I_FSPHI, i, j) = 0.0;
                }
            } else {
                if ((FxL * FxL + Fy * Fy) > 0.0)
                    P2(aux, maux, I_TAUDIR_X, i, j) = -(dx * FxL / sqrt(FxL * FxL + Fy * Fy));
                else
                    P2(aux, maux, I_TAUDIR_X, i, j) = dx * 1.0;
                if ((Fx * Fx + FyL * FyL) > 0.0)
                    P2(aux, maux, I_TAUDIR_Y, i, j) = -(dy * FyL / sqrt(Fx * Fx + FyL * FyL));
                else
                    P2(aux, maux, I_TAUDIR_Y, i, j) = dy * 1.0;
                double tx = P2(aux, maux, I_TAUDIR_X, i, j), ty = P2(aux, maux, I_TAUDIR_Y, i, j);
                if ((ty * ty + tx * tx) > 0.0) P2(aux, maux, I_FSPHI, i, j) = 1.0;
                else P2(aux, maux, I_FSPHI, i, j) = 0.0;
            }
        }
}

/* b4step2: b4step2.f90:7-113 (bed_normal = 0, aux already finalised) */
void oracle_b4step2(oracle_params* P, int mx, int my, int mbc, int maux, double dx, double dy,
                    double* q, double* aux)
{
    const double gz = P->grav;
    for (int j = 1 - mbc; j <= my + mbc; j++)
        for (int i = 1 - mbc; i <= mx + mbc; i++) {
            double u, v, m, chi, rho;
            qfix(P, &P2(q, MEQN, I_H, i, j), &P2(q, MEQN, I_HU, i, j), &P2(q, MEQN, I_HV, i, j),
                 &P2(q, MEQN, I_HM, i, j), &P2(q, MEQN, I_PB, i, j), &P2(q, MEQN, I_HCHI, i, j),
                 &u, &v, &m, &chi, &rho, gz);
            P2(q, MEQN, I_BDIF, i, j) = dmax(P2(q, MEQN, I_BDIF, i, j), 0.0);
        }
    for (int j = 1 - mbc; j <= my + mbc; j++)
        for (int i = 1 - mbc; i <= mx + mbc; i++) {
            if (P2(q, MEQN, 1, i, j) > 0.0) {
                double a = P2(q, MEQN, 2, i, j), b = P2(q, MEQN, 3, i, j);
                double s = sqrt(a * a + b * b) / P2(q, MEQN, 1, i, j);
                if (s > P->speed_limit) {
                    double sratio = P->speed_limit / s;
                    P2(q, MEQN, 2, i, j) = P2(q, MEQN, 2, i, j) * sratio;
                    P2(q, MEQN, 3, i, j) = P2(q, MEQN, 3, i, j) * sratio;
                }
            }
        }
    oracle_calc_taudir(P, mx, my, mbc, maux, dx, dy, q, aux);
}

/* step2: GeoClaw step2.f90 (Clawpack 5; not in the reference tree).
 * x-sweeps over rows j = 0..my+1, then y-sweeps over columns i = 0..mx+1,
 * accumulating into fm/fp/gm/gp in that order. */
void oracle_step2(const oracle_params* P, int mx, int my, int mbc, int maux, const double* qold,
                  const double* aux, double dx, double dy, double dt, double* cflgrid, double* fm,
                  double* fp, double* gm, double* gp)
{
    const int maxm = mx > my ? mx : my;
    const int n = maxm + 2 * mbc;
    double* q1d = (double*)malloc(sizeof(double) * MEQN * n);
    double* aux1 = (double*)malloc(sizeof(double) * maux * n);
    double* aux2 = (double*)malloc(sizeof(double) * maux * n);
    double* aux3 = (double*)malloc(sizeof(double) * maux * n);
    double* dtdx1d = (double*)malloc(sizeof(double) * n);
    double* faddm = (double*)malloc(sizeof(double) * MEQN * n);
    double* faddp = (double*)malloc(sizeof(double) * MEQN * n);
    double* gaddm = (double*)malloc(sizeof(double) * MEQN * n * 2);
    double* gaddp = (double*)malloc(sizeof(double) * MEQN * n * 2);
    const size_t ncell = (size_t)(mx + 2 * mbc) * (my + 2 * mbc);
    *cflgrid = 0.0;
    const double dtdx = dt / dx, dtdy = dt / dy;
    memset(fm, 0, sizeof(double) * MEQN * ncell);
    memset(fp, 0, sizeof(double) * MEQN * ncell);
    memset(gm, 0, sizeof(double) * MEQN * ncell);
    memset(gp, 0, sizeof(double) * MEQN * ncell);
#define GA(arr, m, i, side) (arr)[((m) - 1) + MEQN * (((i) + mbc - 1) + n * ((side) - 1))]

    for (int j = 0; j <= my + 1; j++) {
        for (int i = 1 - mbc; i <= mx + mbc; i++) {
            for (int m = 1; m <= MEQN; m++) S1(q1d, MEQN, m, i) = P2(qold, MEQN, m, i, j);
            for (int m = 1; m <= maux; m++) {
                S1(aux1, maux, m, i) = P2(aux, maux, m, i, j - 1);
                S1(aux2, maux, m, i) = P2(aux, maux, m, i, j);
                S1(aux3, maux, m, i) = P2(aux, maux, m, i, j + 1);
            }
        }
        for (int k = 0; k < n; k++) dtdx1d[k] = dtdx;
        double cfl1d;
        oracle_flux2(P, 1, maxm, mx, mbc, maux, q1d, dtdx1d, aux1, aux2, aux3, faddm, faddp, gaddm,
                     gaddp, &cfl1d);
        *cflgrid = dmax(*cflgrid, cfl1d);
        for (int i = 1; i <= mx + 1; i++)
            for (int m = 1; m <= MEQN; m++) {
                P2(fm, MEQN, m, i, j) = P2(fm, MEQN, m, i, j) + S1(faddm, MEQN, m, i);
                P2(fp, MEQN, m, i, j) = P2(fp, MEQN, m, i, j) + S1(faddp, MEQN, m, i);
                P2(gm, MEQN, m, i, j) = P2(gm, MEQN, m, i, j) + GA(gaddm, m, i, 1);
                P2(gp, MEQN, m, i, j) = P2(gp, MEQN, m, i, j) + GA(gaddp, m, i, 1);
                P2(gm, MEQN, m, i, j + 1) = P2(gm, MEQN, m, i, j + 1) + GA(gaddm, m, i, 2);
                P2(gp, MEQN, m, i, j + 1) = P2(gp, MEQN, m, i, j + 1) + GA(gaddp, m, i, 2);
            }
    }

    for (int i = 0; i <= mx + 1; i++) {
        for (int j = 1 - mbc; j <= my + mbc; j++) {
            for (int m = 1; m <= MEQN; m++) S1(q1d, MEQN, m, j) = P2(qold, MEQN, m, i, j);
            for (int m = 1; m <= maux; m++) {
                S1(aux1, maux, m, j) = P2(aux, maux, m, i - 1, j);
                S1(aux2, maux, m, j) = P2(aux, maux, m, i, j);
                S1(aux3, maux, m, j) = P2(aux, maux, m, i + 1, j);
            }
        }
        for (int k = 0; k < n; k++) dtdx1d[k] = dtdy;
        double cfl1d;
        oracle_flux2(P, 2, maxm, my, mbc, maux, q1d, dtdx1d, aux1, aux2, aux3, faddm, faddp, gaddm,
                     gaddp, &cfl1d);
        *cflgrid = dmax(*cflgrid, cfl1d);
        for (int j = 1; j <= my + 1; j++)
            for (int m = 1; m <= MEQN; m++) {
                P2(gm, MEQN, m, i, j) = P2(gm, MEQN, m, i, j) + S1(faddm, MEQN, m, j);
                P2(gp, MEQN, m, i, j) = P2(gp, MEQN, m, i, j) + S1(faddp, MEQN, m, j);
                P2(fm, MEQN, m, i, j) = P2(fm, MEQN, m, i, j) + GA(gaddm, m, j, 1);
                P2(fp, MEQN, m, i, j) = P2(fp, MEQN, m, i, j) + GA(gaddp, m, j, 1);
                P2(fm, MEQN, m, i + 1, j) = P2(fm, MEQN, m, i + 1, j) + GA(gaddm, m, j, 2);
                P2(fp, MEQN, m, i + 1, j) = P2(fp, MEQN, m, i + 1, j) + GA(gaddp, m, j, 2);
            }
    }
#undef GA
    free(q1d); free(aux1); free(aux2); free(aux3); free(dtdx1d);
    free(faddm); free(faddp); free(gaddm); free(gaddp);
}

/* mp_update_relax_Dclaw4: mp_update.f90:710-789 */
static void mp_update_relax_Dclaw4(const oracle_params* P, double dt, double* h, double* u,
                                   double* v, double* m, double* p, double* chi, double rhoh,
                                   double gz)
{
    const double rho_f = P->rho_f, rho_s = P->rho_s, mu = P->mu;
    double rho, kperm, alphainv, m_eq, tanpsi, tau, hu, hv, hm, hchi;
    setvars(P, *h, *u, *v, *m, *p, *chi, gz, &rho, &kperm, &alphainv, &m_eq, &tanpsi, &tau);
    hu = *h * *u;
    hv = *h * *v;
    hm = *h * *m;
    hchi = *h * *chi;
    double vnorm = sqrt(hu * hu + hv * hv) / *h;
    *p = *p - dt * 3.0 * vnorm * tanpsi * alphainv / *h;
    qfix(P, h, &hu, &hv, &hm, p, &hchi, u, v, m, chi, &rho, gz);

    double zeta = 3.0 * alphainv / (*h * 2.0) + (rho - rho_f) * rho_f * gz / (4.0 * rho);
    double krate = -(zeta * 2.0 * kperm / (*h * dmax(mu, 1e-16)));
    double p_eq = rho_f * *h * gz;
    *p = p_eq + (*p - p_eq) * D_EXP(krate * dt);

    double p_exc = *p - rho_f * gz * *h;
    qfix_cmass(P, h, m, p, &rho, &p_exc, &hu, &hv, &hm, *u, *v, rhoh, gz);
    double D = -(2.0 * (kperm / (mu * *h)) * p_exc);
    krate = D * (rho - rho_f) / rhoh;
    hm = hm * D_EXP(-(dt * D * rho_f / (rhoh)));

    if (P->src2method == 0) {
        *h = *h + *h * krate * dt;
        hchi = *h * *chi;
        hu = hu * D_EXP(dt * krate);
        hv = hv * D_EXP(dt * krate);
        qfix(P, h, &hu, &hv, &hm, p, &hchi, u, v, m, chi, &rho, gz);
    } else if (P->src2method == 1) {
        *h = (rhoh - hm * (rho_s - rho_f)) / rho_f;
        hu = *h * *u;
        hv = *h * *v;
        hchi = *h * *chi;
        qfix(P, h, &hu, &hv, &hm, p, &hchi, u, v, m, chi, &rho, gz);
    }
}

/* mp_update_FE_4quad: mp_update.f90:38-610 (src2method 2).  One substep
 * of the quadrant-based explicit integration of (m, p_exc) on the rho h =
 * const manifold; returns the substep dtk taken out of dt (the time still to
 * integrate).  The branch for vnorm == 0 or m == m_eq (:226-254) reads dtm
 * before any assignment, so its result is not defined by the source: here it
 * sets p to a quiet NaN (check4nans then stops the run, as it would for any
 * NaN) and ends the substeps (dtk = dt).  Quirks kept: region 1B integrates
 * p over dtp but m and the time over dtm (:360-382); region 3A m over dtm,
 * p and the time over dtp (:490-509); case 4 reports dtk = dtp although it
 * integrated over min(dtr, dtp, dtm) (:589-604). */
static double mp_update_FE_4quad(const oracle_params* P, double dt, double* h, double u,
                                 double v, double* m, double* p, double chi, double rhoh,
                                 double gz)
{
    const double rho_f = P->rho_f, rho_s = P->rho_s, mu = P->mu;
    double rho, kperm, alphainv, m_eq, tanpsi, tau, hu, hv, hm, p_exc = 0.0;
    double dtk, dtr, dtp, dtm, dts = 0.0, p_exc_ave;
    const double vnorm = sqrt(u * u + v * v);
    setvars(P, *h, u, v, *m, *p, chi, gz, &rho, &kperm, &alphainv, &m_eq, &tanpsi, &tau);
    const double h0 = *h, m_0 = *m;
    const double rho0 = m_0 * (rho_s - rho_f) + rho_f;
    const double p0 = *p;
    const double p_eq0 = rho_f * gz * h0;
    const double p_exc0 = p0 - p_eq0;
    dtk = 0.0;
    dtr = dt;
    /* :188-190 */
    const double km0 = ((2.0 * kperm * (rho0 * rho0)) / (mu * (rhoh * rhoh)));
    const double c_d0 = -(3.0 * vnorm * (alphainv * rho0 / (rhoh)) * tanpsi);
    double kp0 = (kperm / (h0 * mu)) *
                 (3.0 * alphainv * rho0 / rhoh - 1.5 * rho_f * gz * h0 * (rho0 - rho_f) / rhoh);
    kp0 = dmax(kp0, 1.e-3);                                                   /* :208 */
    if (*m == 0.0) {                                                          /* :212-217 */
        p_exc = p_exc0 * D_EXP(-(kp0 * dtr));
        qfix_cmass(P, h, m, p, &rho, &p_exc, &hu, &hv, &hm, u, v, rhoh, gz);
        return dtr;
    }
    if ((fabs(p_exc0 / (rho_f * gz * h0)) < 1.e-12) && (fabs(*m - m_eq) < 1.e-12))
        return dtr;                                                           /* :220-223 */
    if ((vnorm == 0.0) || (*m == m_eq)) {                                     /* :226-254 */
        *p = NAN;   /* dtm is read unset by the reference: undefined result */
        return dtr;
    }
    /* critical point (m_c, p_eq_c): fixed-point iteration, :257-293 */
    double m_c = P->m_crit;
    double rho_c = m_c * (rho_s - rho_f) + rho_f;
    double h_c = rhoh / rho_c;
    double p_eq_c = rho_f * gz * h_c;
    const double shear = 2.0 * vnorm / h0;
    const double convtol = 1.e-16;
    const int itermax = 100;
    double normc = 0.0;
    int iter;
    for (iter = 1; iter <= itermax; iter++) {
        const double m_c0 = m_c, p_eq_c0 = p_eq_c;
        rho_c = m_c * (rho_s - rho_f) + rho_f;
        h_c = rhoh / rho_c;
        p_eq_c = rho_f * gz * h_c;
        const double sig_c = dmax(rhoh * gz - p_eq_c, 0.0);
        const double sd = shear * P->delta;
        const double Nd = rho_s * (sd * sd) + sig_c;
        const double Nn = mu * shear;
        m_c = P->m_crit * (sqrt(Nd) / (sqrt(Nd) + sqrt(Nn)));
        const double a = 1.e2 * (m_c - m_c0), b = 1.e-3 * (p_eq_c - p_eq_c0);
        normc = a * a + b * b;
        if (normc < convtol) break;
    }
    /* a Fortran DO loop leaves iter = itermax + 1 when it runs to the end */
    if ((iter >= itermax) && (normc > 1.0)) {
        m_c = P->m_crit;
        rho_c = m_c * (rho_s - rho_f) + rho_f;
        h_c = rhoh / rho_c;
        p_eq_c = rho_f * gz * h_c;
    } else {
        rho_c = m_c * (rho_s - rho_f) + rho_f;
        h_c = rhoh / rho_c;
        p_eq_c = rho_f * gz * h_c;
    }
    int quad0 = 0;                                                            /* :297-305 */
    if ((*m < m_c) && (*p >= p_eq_c)) quad0 = 1;
    else if ((*m >= m_c) && (*p > p_eq_c)) quad0 = 2;
    else if ((*m > m_c) && (*p <= p_eq_c)) quad0 = 3;
    else if ((*m <= m_c) && (*p < p_eq_c)) quad0 = 4;

    switch (quad0) {
    case 1:                                                                   /* :311-422 */
        if (p_exc0 < -(1.e-3 * rhoh * gz)) {                                  /* region 1A */
            if (c_d0 > 0.0) dtp = -(1.0 / kp0) * D_LOG(c_d0 / (c_d0 - kp0 * p_exc0));
            else dtp = dtr;
            dtm = D_LOG(0.98) / (kp0 * p_exc0);
            dts = dmin(dmin(dtr, dtp), dtm);
            dts = dmax(dts, 0.0);
            p_exc = (1.0 / kp0) * ((kp0 * p_exc0 - c_d0) * D_EXP(-(kp0 * dts)) + c_d0);
            *m = m_0 * D_EXP(dts * km0 * p_exc);
        } else if ((-(kp0 * p_exc0) + c_d0) > 0.0) {                          /* region 1B */
            if ((c_d0 / kp0) > (rhoh * gz - p_eq_c))
                dtp = dmin(dtr, -(1.0 / kp0) * D_LOG((-(kp0 * rhoh * gz) + kp0 * p_eq_c + c_d0) /
                                                     (-(kp0 * p_exc0) + c_d0)));
            else
                dtp = dtr;
            p_exc = (1.0 / kp0) * ((kp0 * p_exc0 - c_d0) * D_EXP(-(kp0 * dtp)) + c_d0);
            p_exc_ave = 0.5 * (p_exc + p_exc0);
            if (p_exc_ave <= 0.0) {
                dtm = dtr;
                *m = m_0 * D_EXP(km0 * 0.5 * (p_exc + p_exc0) * dtm);
            } else {
                if ((km0 * p_exc_ave * m_0) < 1.e-99) dtm = dtp;
                else dtm = dmin(dtp, (m_eq - m_0) / (km0 * p_exc_ave * m_0));
            }
            *m = m_0 + km0 * m_0 * p_exc_ave * dtm;
            dts = dtm;
        } else {                                                              /* region 1C */
            if (c_d0 > 0.0) {
                dtp = dtr;
                if ((km0 * p_exc0 * m_0) > 1.e-99) dtm = (m_c - m_0) / (km0 * p_exc0 * m_0);
                else dtm = dtp;
                dts = dmin(dtm, dtp);
                dts = dmax(dts, 0.0);
                *m = m_0 + km0 * m_0 * p_exc0 * dts;
                p_exc = (1.0 / kp0) * ((kp0 * p_exc0 - c_d0) * D_EXP(-(kp0 * dts)) + c_d0);
            } else {
                dts = dtr;
                dtp = dtr;
                if ((km0 * p_exc0 * m_0) > 1.e-99) dts = dmin(dtr, (m_c - m_0) / (km0 * p_exc0 * m_0));
                if ((kp0 * p_exc0 - c_d0) != 0.0) {
                    if ((-(kp0 * p_eq0) - c_d0) / (kp0 * p_exc0 - c_d0) > 1.e-99)
                        dtp = dmin(dtr, -(1.0 / kp0) * D_LOG((c_d0) / (c_d0 - kp0 * p_exc0)));
                }
                dts = dmin(dts, dtp);
                dts = dmax(dts, 0.0);
                *m = m_0 + km0 * m_0 * p_exc0 * dts;
                p_exc = (1.0 / kp0) * ((kp0 * p_exc0 - c_d0) * D_EXP(-(kp0 * dts)) + c_d0);
            }
        }
        qfix_cmass(P, h, m, p, &rho, &p_exc, &hu, &hv, &hm, u, v, rhoh, gz);
        dtk = dts;
        break;
    case 2:                                                                   /* :424-476 */
        dtp = dtr;
        if ((c_d0 - kp0 * (p_eq_c - p_eq0)) / (c_d0 - kp0 * p_exc0) > 1.e-99)
            dtp = dmin(dtr, -(1.0 / kp0) * D_LOG((c_d0 - kp0 * (p_eq_c - p_eq0)) / (c_d0 - kp0 * p_exc0)));
        dtm = dtp;
        if (km0 * p_exc0 * m_0 > 1.e-99)
            dtm = dmin(dtp, (dmin(1.0 - m_0, 0.02) / (km0 * p_exc0 * m_0)));
        dts = dmin(dtp, dtm);
        p_exc = (1.0 / kp0) * ((kp0 * p_exc0 - c_d0) * D_EXP(-(kp0 * dts)) + c_d0);
        *m = m_0 + km0 * m_0 * p_exc * dts;
        qfix_cmass(P, h, m, p, &rho, &p_exc, &hu, &hv, &hm, u, v, rhoh, gz);
        dtk = dts;
        break;
    case 3:                                                                   /* :478-571 */
        if (p_exc0 > 1.e-3 * rhoh * gz) {                                     /* region 3A */
            dtp = dtr;
            if ((c_d0 / (-(kp0 * p_exc0) + c_d0)) > 1.e-99)
                dtp = dmin(dtr, -(1.0 / kp0) * D_LOG(c_d0 / (c_d0 - kp0 * p_exc0)));
            dtm = dtp;
            if ((km0 * p_exc0 * m_0) > 1.e-99)
                dtm = dmin(dtp, (dmin(1.0 - m_0, 0.02)) / (km0 * p_exc0 * m_0));
            p_exc = (1.0 / kp0) * ((kp0 * p_exc0 - c_d0) * D_EXP(-(kp0 * dtp)) + c_d0);
            *m = m_0 + km0 * m_0 * 0.5 * (p_exc0 + p_exc) * dtm;
            dts = dtp;
        } else if ((-(kp0 * p_exc0) + c_d0) < 0.0) {                          /* region 3B */
            if (c_d0 / kp0 < -p_eq0) {
                dtp = dtr;
                if (((c_d0 + kp0 * p_eq0) / (-(kp0 * p_exc0) + c_d0)) > 1.e-99)
                    dtp = dmin(dtr, -(1.0 / kp0) * D_LOG((c_d0 + kp0 * p_eq0) / (-(kp0 * p_exc0) + c_d0)));
            } else {
                dtp = dtr;
            }
            p_exc = (1.0 / kp0) * ((kp0 * p_exc0 - c_d0) * D_EXP(-(kp0 * dtp)) + c_d0);
            p_exc_ave = 0.5 * (p_exc + p_exc0);
            if (p_exc_ave > 0.0) {
                dtm = dmin(dtr, (0.02) / (km0 * p_exc_ave * m_0));
            } else {
                const double m_lower = m_eq + dmax(0.0, (1.0 / 3.0) * kp0 * h0 * (-p_exc_ave) /
                                                            (vnorm * alphainv));
                dtm = dmin(dtr, (m_lower - m_0) / (km0 * p_exc_ave * m_0));
            }
            *m = m_0 + km0 * m_0 * p_exc_ave * dtm;
            dts = dtm;
        } else {                                                              /* region 3C */
            if (c_d0 < 0.0) {
                dtp = dtr;
                dtm = dtr;
                if (-(km0 * p_exc0 * m_0) > 1.e-99) dtm = (m_c - m_0) / (km0 * p_exc0 * m_0);
                dts = dmin(dtp, dtm);
                dts = dmax(dts, 0.0);
                p_exc = (1.0 / kp0) * ((kp0 * p_exc0 - c_d0) * D_EXP(-(kp0 * dts)) + c_d0);
                *m = m_0 + km0 * m_0 * p_exc0 * dts;
            } else {
                dtm = dtr;
                dtp = dtr;
                if ((km0 * p_exc0 * m_0) > 1.e-99) dtm = (m_c - m_0) / (km0 * p_exc0 * m_0);
                if ((-c_d0 / (kp0 * p_exc0 - c_d0)) > 1.e-99)
                    dtp = -(1.0 / kp0) * D_LOG(c_d0 / (c_d0 - kp0 * p_exc0));
                dts = dmin(dtp, dtm);
                dts = dmax(dts, 0.0);
                p_exc = (1.0 / kp0) * ((kp0 * p_exc0 - c_d0) * D_EXP(-(kp0 * dts)) + c_d0);
                *m = m_0 + km0 * m_0 * p_exc0 * dts;
            }
        }
        qfix_cmass(P, h, m, p, &rho, &p_exc, &hu, &hv, &hm, u, v, rhoh, gz);
        dtk = dts;
        break;
    case 4:                                                                   /* :573-605 */
        dtp = dtr;
        if ((c_d0 - kp0 * (p_eq_c - p_eq0)) / (c_d0 - kp0 * p_exc0) > 1.e-99)
            dtp = dmin(dtr, -(1.0 / kp0) * D_LOG((c_d0 - kp0 * (p_eq_c - p_eq0)) / (c_d0 - kp0 * p_exc0)));
        dtm = D_LOG(0.98) / (kp0 * p_exc0);
        dts = dmin(dtr, dmin(dtp, dtm));
        p_exc = (1.0 / kp0) * ((kp0 * p_exc0 - c_d0) * D_EXP(-(kp0 * dts)) + c_d0);
        *m = m_0 * D_EXP(dts * km0 * 0.5 * (p_exc0 + p_exc));
        dts = dtp;
        qfix_cmass(P, h, m, p, &rho, &p_exc, &hu, &hv, &hm, u, v, rhoh, gz);
        dtk = dts;
        break;
    default:
        break;
    }
    (void)tau;
    return dtk;
}

/* src2: src2.f90:27-306 (bed_normal = curvature = 0; src2method -1, 0, 1,
 * 2) */
/* ent_dclaw4: entrainment.f90:27-100.  b_x and b_y are arguments of the
 * reference routine that it never reads; they are left out. */
static void ent_dclaw4(const oracle_params* P, double dt, double* h, double u, double v,
                       double* m, double* p, double* rho, double* hchi, double gz, double tau,
                       double* b_eroded, double b_remaining)
{
    double coeff, vlow, vnorm, rhoe, beta, visc, gamma, t1bot, t2top, prat, dh, hm, hrho;
    if (P->friction_forcing) coeff = P->manning_coefficient;      /* :44-48 */
    else coeff = 0.0;
    vnorm = sqrt(u * u + v * v);
    vlow = 0.1;
    if (vnorm > vlow) {
        if (b_remaining > 0.0) {
            rhoe = P->me * P->rho_s + (1.0 - P->me) * P->rho_f;   /* :62 */
            beta = dmax(1.0 - *m, 0.1);                            /* :66 */
            visc = beta * vnorm * 2.0 * P->mu * (1.0 - *m) / (D_TANH(*h + 1.e-2));
            gamma = *rho * beta * (vnorm * vnorm) * (beta * gz * (coeff * coeff)) /
                    (D_POW(D_TANH(*h + 1.e-2), 1.0 / 3.0));        /* :69 */
            t1bot = visc + gamma + tau;
            t2top = dmin(t1bot, (1.0 - beta * P->entrainment_rate) * tau);
            prat = *p / (*rho * *h);                               /* :75 */
            dh = P->entrainment_rate * dt * (t1bot - t2top) / (rhoe * beta * vnorm);
            dh = dmin(dh, b_remaining);                            /* :80 */
            *h = *h + dh;
            hm = *h * *m + dh * P->me;
            *m = hm / *h;
            *hchi = *hchi + dh * P->chie;                          /* :88 */
            hrho = hm * P->rho_s + (*h - hm) * P->rho_f;
            *rho = hrho / *h;
            *b_eroded = *b_eroded + dh;                            /* :93 */
            *p = prat * hrho;
        }
    }
}

void oracle_src2(const oracle_params* P, int mx, int my, int mbc, int maux, double dx,
                 double dy, double* q, const double* aux, double dt)
{
    const double drytol = P->dry_tolerance, rho_s = P->rho_s, rho_f = P->rho_f, mu = P->mu;
    double coeffmanning = P->friction_forcing ? P->manning_coefficient : 0.0;
    /* gz is set once (src2.f90:87) and, with curvature, each wet cell adds its
     * curvature acceleration to it (:156) without a reset: a cell's gravity
     * includes every earlier cell's in loop order (j outer, i inner) */
    double gz = P->grav;
    const double gx = 0.0;
    for (int j = 1; j <= my; j++)
        for (int i = 1; i <= mx; i++) {
            double h = P2(q, MEQN, I_H, i, j);
            if (h <= drytol) {
                for (int m = 1; m <= 6; m++) P2(q, MEQN, m, i, j) = 0.0;
                continue;
            }
            double hu = P2(q, MEQN, I_HU, i, j), hv = P2(q, MEQN, I_HV, i, j);
            double hm = P2(q, MEQN, I_HM, i, j), p = P2(q, MEQN, I_PB, i, j);
            double hchi = P2(q, MEQN, I_HCHI, i, j);
            double rhoh = hm * rho_s + (h - hm) * rho_f;
            if (hm < 0.0) continue;
            if (P->curvature == 1) {                                   /* :133-156 */
#define BD(ii, jj) (P2(aux, maux, 1, ii, jj) - P2(q, MEQN, I_BDIF, ii, jj))
                const double b = BD(i, j), bL = BD(i - 1, j), bR = BD(i + 1, j);
                const double bT = BD(i, j + 1), bB = BD(i, j - 1);
                const double bTR = BD(i + 1, j + 1), bTL = BD(i - 1, j + 1);
                const double bBR = BD(i + 1, j - 1), bBL = BD(i - 1, j - 1);
#undef BD
                const double b_xx = (bR - 2.0 * b + bL) / (dx * dx);
                const double b_yy = (bT - 2.0 * b + bB) / (dy * dy);
                const double b_xy = ((bTR - bTL) - (bBR - bBL)) / (4.0 * dx * dy);
                const double uc = hu / h, vc = hv / h;
                const double dtheta = -((P2(aux, maux, I_THETA, i + 1, j) -
                                         P2(aux, maux, I_THETA, i - 1, j)) / (2.0 * dx));
                const double gacc = dmax(uc * uc * b_xx + vc * vc * b_yy + 2.0 * uc * vc * b_xy +
                                         uc * uc * dtheta, 0.0);
                gz = gz + gacc;
            }

            double u, v, m, chi, rho, kperm, alphainv, m_eq, tanpsi, tau;
            qfix(P, &h, &hu, &hv, &hm, &p, &hchi, &u, &v, &m, &chi, &rho, gz);
            setvars(P, h, u, v, m, p, chi, gz, &rho, &kperm, &alphainv, &m_eq, &tanpsi, &tau);

            double hvnorm0 = sqrt(hu * hu + hv * hv);
            double vnorm = hvnorm0 / h;
            if (hvnorm0 > 0.0) {
                vnorm = dmax(0.0, vnorm - dt * tau / rhoh);
                vnorm = vnorm * D_EXP(-((1.0 - m) * 2.0 * mu * dt / (h * rhoh)));
                double hvnorm = h * vnorm;
                hu = hvnorm * hu / hvnorm0 + gx * h * dt;
                hv = hvnorm * hv / hvnorm0;
                u = hu / h;
                v = hv / h;
                vnorm = sqrt(u * u + v * v);
            }
            setvars(P, h, u, v, m, p, chi, gz, &rho, &kperm, &alphainv, &m_eq, &tanpsi, &tau);

            if (P->src2method == -1) {
                hu = h * u; hv = h * v; hm = h * m; hchi = h * chi;
            } else if (P->src2method == 2) {
                /* src2.f90:209-226: substeps until dt is used up */
                double dtremaining = dt;
                const int itercountmax = 100;
                int itercount = 0;
                while (dtremaining > 1.e-99) {
                    const double dtk = mp_update_FE_4quad(P, dtremaining, &h, u, v, &m, &p, chi,
                                                          rhoh, gz);
                    dtremaining = dtremaining - dtk;
                    itercount = itercount + 1;
                    if (dtk == 0.0) break;
                    if (h < drytol) break;
                    if (itercount >= itercountmax) break;
                }
                hu = h * u; hv = h * v; hm = h * m; hchi = h * chi;
                qfix(P, &h, &hu, &hv, &hm, &p, &hchi, &u, &v, &m, &chi, &rho, gz);
            } else {
                mp_update_relax_Dclaw4(P, dt, &h, &u, &v, &m, &p, &chi, rhoh, gz);
                hu = h * u; hv = h * v; hm = h * m; hchi = h * chi;
            }

            if (h <= drytol) {
                for (int mm = 1; mm <= 6; mm++) P2(q, MEQN, mm, i, j) = 0.0;
                continue;
            }
            setvars(P, h, u, v, m, p, chi, gz, &rho, &kperm, &alphainv, &m_eq, &tanpsi, &tau);

            /* mass entrainment, src2.f90:253-263 (aux(i_ent) = aux(7)) */
            if (P->entrainment == 1) {
                double b_eroded = P2(q, MEQN, I_BDIF, i, j);
                double b_remaining = P2(aux, maux, 7, i, j) - b_eroded;
                if (P->entrainment_method == 0) {
                    ent_dclaw4(P, dt, &h, u, v, &m, &p, &rho, &hchi, gz, tau, &b_eroded,
                               b_remaining);
                    P2(q, MEQN, I_BDIF, i, j) = b_eroded;
                }
            }

            if (P->friction_forcing && coeffmanning > 0.0) {
                if (h <= P->friction_depth) {
                    if (P->dd_manning)
                        coeffmanning = P->manning_max + (P->manning_coefficient - P->manning_max)
                                       * 0.5 * (1.0 + D_TANH(300.0 * (h - 0.05)));
                    double beta = 1.0;
                    double gamma = beta * sqrt(hu * hu + hv * hv) * (gz * (coeffmanning * coeffmanning))
                                   / D_POW(h, 7.0 / 3.0);
                    double dgamma = 1.0 + dt * gamma;
                    hu = hu / dgamma;
                    hv = hv / dgamma;
                    qfix(P, &h, &hu, &hv, &hm, &p, &hchi, &u, &v, &m, &chi, &rho, gz);
                }
            }
            P2(q, MEQN, I_H, i, j) = h;
            P2(q, MEQN, I_HU, i, j) = hu;
            P2(q, MEQN, I_HV, i, j) = hv;
            P2(q, MEQN, I_HM, i, j) = hm;
            P2(q, MEQN, I_PB, i, j) = p;
            P2(q, MEQN, I_HCHI, i, j) = hchi;
        }
}

/* stepgrid: GeoClaw stepgrid.f (Clawpack 5; not in the reference tree):
 * b4step2, step2, conservative update of interior cells, Godunov-split src2. */
double oracle_stepgrid(oracle_params* P, int mx, int my, int mbc, int maux, double dx,
                       double dy, double dt, double* q, double* aux, double* work)
{
    const size_t ncell = (size_t)(mx + 2 * mbc) * (my + 2 * mbc);
    double* fm = work;
    double* fp = work + MEQN * ncell;
    double* gm = work + 2 * MEQN * ncell;
    double* gp = work + 3 * MEQN * ncell;
    double cflgrid;
    oracle_b4step2(P, mx, my, mbc, maux, dx, dy, q, aux);
    oracle_step2(P, mx, my, mbc, maux, q, aux, dx, dy, dt, &cflgrid, fm, fp, gm, gp);
    const double dtdx = dt / dx, dtdy = dt / dy;
    for (int j = 1; j <= my; j++)
        for (int i = 1; i <= mx; i++)
            for (int m = 1; m <= MEQN; m++)
                P2(q, MEQN, m, i, j) = P2(q, MEQN, m, i, j)
                    - dtdx * (P2(fm, MEQN, m, i + 1, j) - P2(fp, MEQN, m, i, j))
                    - dtdy * (P2(gm, MEQN, m, i, j + 1) - P2(gp, MEQN, m, i, j));
    oracle_src2(P, mx, my, mbc, maux, dx, dy, q, aux, dt);
    return cflgrid;
}

/* Physical boundary conditions (GeoClaw bc2amr; not in the reference tree):
 * extrap copies the nearest interior cell, wall mirrors with the normal
 * momentum negated.  x sides first for interior rows, then y sides across
 * the full width (corner values are the same in either order for these
 * two rules). */
void oracle_bc_fill(int mx, int my, int mbc, int meqn, double* q, const int* bc)
{
#define QQ(m, i, j) q[((m) - 1) + (size_t)meqn * ((size_t)((i) + mbc - 1) + (size_t)((j) + mbc - 1) * (size_t)(mx + 2 * mbc))]
    for (int j = 1; j <= my; j++)
        for (int k = 1; k <= mbc; k++)
            for (int m = 1; m <= meqn; m++) {
                if (bc[0] == 3) QQ(m, 1 - k, j) = (m == 2 ? -1.0 : 1.0) * QQ(m, k, j);
                else QQ(m, 1 - k, j) = QQ(m, 1, j);
                if (bc[1] == 3) QQ(m, mx + k, j) = (m == 2 ? -1.0 : 1.0) * QQ(m, mx + 1 - k, j);
                else QQ(m, mx + k, j) = QQ(m, mx, j);
            }
    for (int i = 1 - mbc; i <= mx + mbc; i++)
        for (int k = 1; k <= mbc; k++)
            for (int m = 1; m <= meqn; m++) {
                if (bc[2] == 3) QQ(m, i, 1 - k) = (m == 3 ? -1.0 : 1.0) * QQ(m, i, k);
                else QQ(m, i, 1 - k) = QQ(m, i, 1);
                if (bc[3] == 3) QQ(m, i, my + k) = (m == 3 ? -1.0 : 1.0) * QQ(m, i, my + 1 - k);
                else QQ(m, i, my + k) = QQ(m, i, my);
            }
#undef QQ
}

int oracle_run(oracle_params* P, int mx, int my, int mbc, int maux, double dx, double dy,
               const int* bc, double* q, double* aux, int nsteps, double* dt_inout,
               double cfl_desired, double dt_max, double* cfl_hist, double* dt_hist,
               int nthreads)
{
    (void)nthreads;   /* one patch runs on one thread; oracle_run_patches spreads patches */
    const size_t ncell = (size_t)(mx + 2 * mbc) * (my + 2 * mbc);
    double* work = (double*)malloc(sizeof(double) * 4 * MEQN * ncell);
    double dt = *dt_inout;
    int n;
    for (n = 0; n < nsteps; n++) {
        oracle_bc_fill(mx, my, mbc, MEQN, q, bc);
        double cfl = oracle_stepgrid(P, mx, my, mbc, maux, dx, dy, dt, q, aux, work);
        if (dt_hist) dt_hist[n] = dt;
        if (cfl_hist) cfl_hist[n] = cfl;
        double dtnew = (cfl > 0.0) ? dt * cfl_desired / cfl : INFINITY;
        dt = dmin(dtnew, dt_max);
    }
    *dt_inout = dt;
    free(work);
    return n;
}

/* Many patches, each with its own parameter record and dt rule, advanced
 * nsteps each: the reference's OpenMP over patches (AMRClaw advanc runs
 * stepgrid for the grids of a level in an OpenMP loop with a dynamic
 * schedule, update.f90:50-60 and amr2.f90:592 name the same structure).
 * Built with OpenMP only in the timing variant of the library; elsewhere the
 * loop runs on one thread.  Patches are independent here (ensemble members
 * or patches with fixed ghost rules), so no ghost exchange is needed. */
int oracle_run_patches(oracle_params* P, int npatch, const int* mx, const int* my, int mbc,
                       const int* maux, const double* dx, const double* dy, const int* bc,
                       double** q, double** aux, int nsteps, double* dt_inout,
                       double cfl_desired, double dt_max, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int k = 0; k < npatch; k++)
        oracle_run(&P[k], mx[k], my[k], mbc, maux[k], dx[k], dy[k], bc, q[k], aux[k], nsteps,
                   &dt_inout[k], cfl_desired, dt_max, NULL, NULL, 1);
    return npatch;
}

int oracle_openmp_threads(void)
{
#ifdef _OPENMP
    return 1 << 20;
#else
    return 1;
#endif
}

void oracle_dt_rule(double dt_used, double cfl, double t_before, double cfl_desired,
                    double dt_max, double tout, double* t_after, double* dt_next, int* hit)
{
    const double t = t_before + dt_used;
    double dtn = (cfl > 0.0) ? dt_used * cfl_desired / cfl : INFINITY;
    dtn = dmin(dtn, dt_max);
    *hit = 0;
    if (t < tout && t + 1.001 * dtn >= tout) {   /* tick.f:135-144 */
        dtn = tout - t;
        *hit = 1;
    }
    *t_after = t;
    *dt_next = dtn;
}

void oracle_gauges(int mx, int my, int mbc, int maux, const double* q, const double* aux,
                   double xlower, double ylower, double dx, double dy, double dry_tolerance,
                   const double* xg, const double* yg, int ngauges, double* out)
{
    (void)my;
    for (int g = 0; g < ngauges; g++) {
        const double fx = (xg[g] - xlower) / dx, fy = (yg[g] - ylower) / dy;
        const int ic = (int)floor(fx) + 1, jc = (int)floor(fy) + 1;   /* containing cell */
        const int i0 = (int)floor(fx + 0.5), j0 = (int)floor(fy + 0.5);  /* lower-left centre */
        const double xo = fx + 0.5 - (double)i0, yo = fy + 0.5 - (double)j0;
        int wet = 1;
        for (int dj = 0; dj < 2; dj++)
            for (int di = 0; di < 2; di++)
                if (!(P2(q, MEQN, 1, i0 + di, j0 + dj) > dry_tolerance)) wet = 0;
        double* o = out + 8 * g;
        if (wet) {
            const double w00 = (1.0 - xo) * (1.0 - yo), w10 = xo * (1.0 - yo);
            const double w01 = (1.0 - xo) * yo, w11 = xo * yo;
            for (int m = 1; m <= MEQN; m++)
                o[m - 1] = ((w00 * P2(q, MEQN, m, i0, j0) + w10 * P2(q, MEQN, m, i0 + 1, j0)) +
                            w01 * P2(q, MEQN, m, i0, j0 + 1)) +
                           w11 * P2(q, MEQN, m, i0 + 1, j0 + 1);
            double topo = 0.0;
            for (int dj = 0; dj < 2; dj++)
                for (int di = 0; di < 2; di++) {
                    const double w = (di ? xo : 1.0 - xo) * (dj ? yo : 1.0 - yo);
                    topo = topo + w * (P2(aux, maux, 1, i0 + di, j0 + dj) -
                                       P2(q, MEQN, MEQN, i0 + di, j0 + dj));
                }
            o[7] = o[0] + topo;
        } else {
            for (int m = 1; m <= MEQN; m++) o[m - 1] = P2(q, MEQN, m, ic, jc);
            o[7] = o[0] + (P2(aux, maux, 1, ic, jc) - P2(q, MEQN, MEQN, ic, jc));
        }
    }
}

/* ---- point-routine entry points for tests ---- */
void oracle_qfix(const oracle_params* P, double* q6, double gz, double* prim5)
{
    qfix(P, &q6[0], &q6[1], &q6[2], &q6[3], &q6[4], &q6[5], &prim5[0], &prim5[1], &prim5[2],
         &prim5[3], &prim5[4], gz);
}

void oracle_setvars(const oracle_params* P, const double* in7, double* out6)
{
    setvars(P, in7[0], in7[1], in7[2], in7[3], in7[4], in7[5], in7[6], &out6[0], &out6[1],
            &out6[2], &out6[3], &out6[4], &out6[5]);
}

void oracle_riemanntype(double hL, double hR, double uL, double uR, int maxiter, double drytol,
                        double g, double* out3, int* rare2)
{
    int r1, r2;
    riemanntype(hL, hR, uL, uR, &out3[0], &out3[1], &out3[2], &r1, &r2, maxiter, drytol, g);
    rare2[0] = r1;
    rare2[1] = r2;
}

/* in30: hL hR huL huR hvL hvR hmL hmR pL pR bL bR uL uR vL vR mL mR thetaL
 * thetaR taudirL taudirR chiL chiR fsL fsR (26 used) */
void oracle_riemann_solver(const oracle_params* P, int ixy, const double* x, int wallprob,
                           double* sw3, double* fw18)
{
    riemann_dig2_aug_sswave_ez(P, ixy, x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8],
                               x[9], x[10], x[11], x[12], x[13], x[14], x[15], x[16], x[17],
                               x[18], x[19], sw3, fw18, wallprob, x[20], x[21], x[22], x[23],
                               x[24], x[25]);
}

/* ------------------------------------------------------------------------
 * AMR transfers between a coarse patch and a fine patch (coordinate_system 1,
 * mcapa = 0: capa = capac = 1).  Arrays are (7, ni, nj) / (maux, ni, nj),
 * component fastest, indices 0-based over the whole array including ghosts.
 * ------------------------------------------------------------------------ */

/* update.f90:112-203 (fine -> coarse): coarse cells (ci0 + a, cj0 + b),
 * a < ni, b < nj, each the average of the rx x ry fine cells from
 * (fi0 + a rx, fj0 + b ry): depth from the mean surface of the wet fine
 * cells (nwet of them) capped by their mean depth, momenta (components 2..7)
 * of the wet fine cells scaled by hc / hsum, where hsum sums the depth of
 * every fine cell, dry ones included (update.f90:160-164). */
void oracle_amr_average(double* qc, const double* auxc, int mauxc, int mitc, const double* qf,
                        const double* auxf, int mauxf, int mitf, int ci0, int cj0, int ni, int nj,
                        int fi0, int fj0, int rx, int ry, double dry_tolerance)
{
    for (int a = 0; a < ni; a++) {          /* update.f90:124-125: i outer, j inner */
        for (int b = 0; b < nj; b++) {
            const int ic = ci0 + a, jc = cj0 + b;
            const int if0 = fi0 + a * rx, jf0 = fj0 + b * ry;
            const double capac = 1.0;
            const double bc = auxc[(size_t)mauxc * ((size_t)ic + (size_t)mitc * jc)];
            double etasum = 0.0, hsum = 0.0, husum[7], huc[7];
            for (int m = 1; m < 7; m++) husum[m] = 0.0;
            int nwet = 0;
            for (int jco = 0; jco < ry; jco++) {
                for (int ico = 0; ico < rx; ico++) {
                    const size_t cf = (size_t)(if0 + ico) + (size_t)mitf * (jf0 + jco);
                    const double* f = qf + 7 * cf;
                    const double capa = 1.0;
                    const double hf = f[0] * capa;
                    const double bf = auxf[(size_t)mauxf * cf] * capa;
                    double huf[7], etaf;
                    for (int m = 1; m < 7; m++) huf[m] = f[m] * capa;
                    if (f[0] > dry_tolerance) {
                        etaf = hf + bf;
                        nwet = nwet + 1;
                    } else {
                        etaf = 0.0;
                        for (int m = 1; m < 7; m++) huf[m] = 0.0;
                    }
                    hsum = hsum + hf;
                    for (int m = 1; m < 7; m++) husum[m] = husum[m] + huf[m];
                    etasum = etasum + etaf;
                }
            }
            double hc;
            if (nwet > 0) {
                const double etaav = etasum / (double)nwet;
                const double hav = hsum / (double)nwet;
                hc = dmin(hav, dmax(etaav - bc * capac, 0.0));
                const double fac = dmin(hav, hc) / hsum;
                for (int m = 1; m < 7; m++) huc[m] = fac * husum[m];
            } else {
                hc = 0.0;
                for (int m = 1; m < 7; m++) huc[m] = 0.0;
            }
            double* c = qc + 7 * ((size_t)ic + (size_t)mitc * jc);
            c[0] = hc / capac;
            for (int m = 1; m < 7; m++) c[m] = huc[m] / capac;
        }
    }
}

/* filval.f90:190-381 (coarse -> fine, the interpolation loop): valc/auxc is
 * the coarse patch enlarged by one coarse cell all round (mic x mjc, filled
 * beforehand); coarse cell (i, j), 1 <= i <= mic-2, 1 <= j <= mjc-2, fills
 * the fine cells ((i-1) rx + nghost + ico, (j-1) ry + nghost + jco),
 * 0 <= ico < rx, 0 <= jco < ry, of val (mitot x mjtot, nghost = 2).  Depth:
 * the surface (h + b, or sea_level for a dry coarse cell) limited-linear
 * (minmod) interpolated, minus the fine bed, clipped at 0, only in fine cells
 * with needs_set != 0 (the reference's NEEDS_TO_BE_SET flag).  Components
 * 2..7: limited-linear interpolation kept when every fine "velocity"
 * hvf / h stays within the range of the coarse cell and its wet neighbours,
 * else (or if a set fine cell came out dry) the coarse value spread in
 * proportion to the fine depth, clamped to that range for components 4..7
 * (filval.f90:289-381); this part runs over every fine cell of the block
 * whether or not it was set.  use_force_dry and variable_eta_init are off. */
void oracle_amr_interp(const double* valc, const double* auxc, int mauxc, int mic, int mjc,
                       double* val, const double* auxf, int mauxf, int mitot,
                       const unsigned char* needs_set, int rx, int ry, double dry_tolerance,
                       double sea_level, double rho_f, double grav)
{
    const int nghost = 2;
#define VC(m, i, j) valc[(m) + 7 * ((size_t)(i) + (size_t)mic * (j))]
#define BCX(i, j) auxc[(size_t)mauxc * ((size_t)(i) + (size_t)mic * (j))]
#define VF(m, i, j) val[(m) + 7 * ((size_t)(i) + (size_t)mitot * (j))]
    for (int j = 1; j <= mjc - 2; j++) {        /* filval.f90:190-191: j outer, i inner */
        for (int i = 1; i <= mic - 2; i++) {
            int fineflag1 = 0;
            double cv[3];
            for (int ii = -1; ii <= 1; ii++) {
                cv[1 + ii] = VC(0, i + ii, j) + BCX(i + ii, j);
                if (VC(0, i + ii, j) <= dry_tolerance) cv[1 + ii] = sea_level;
            }
            double s1p = cv[2] - cv[1], s1m = cv[1] - cv[0];
            double slopex = dmin(fabs(s1p), fabs(s1m)) * dsign(1.0, cv[2] - cv[0]);
            if (s1m * s1p <= 0.0) slopex = 0.0;
            for (int jj = -1; jj <= 1; jj++) {
                cv[1 + jj] = VC(0, i, j + jj) + BCX(i, j + jj);
                if (VC(0, i, j + jj) <= dry_tolerance) cv[1 + jj] = sea_level;
            }
            s1p = cv[2] - cv[1];
            s1m = cv[1] - cv[0];
            double slopey = dmin(fabs(s1p), fabs(s1m)) * dsign(1.0, cv[2] - cv[0]);
            if (s1m * s1p <= 0.0) slopey = 0.0;
            double finemass = 0.0;
            for (int jco = 1; jco <= ry; jco++) {
                for (int ico = 1; ico <= rx; ico++) {
                    const double yoff = ((double)jco - 0.5) / (double)ry - 0.5;
                    const double xoff = ((double)ico - 0.5) / (double)rx - 0.5;
                    const int jf = (j - 1) * ry + nghost + jco - 1;
                    const int ifn = (i - 1) * rx + nghost + ico - 1;
                    if (!needs_set[(size_t)ifn + (size_t)mitot * jf]) continue;
                    double v = (cv[1] + xoff * slopex) + yoff * slopey;
                    v = dmax(0.0, v - auxf[(size_t)mauxf * ((size_t)ifn + (size_t)mitot * jf)]);
                    VF(0, ifn, jf) = v;
                    finemass = finemass + v;
                    if (v <= dry_tolerance) {
                        fineflag1 = 1;
                        for (int m = 1; m < 7; m++) VF(m, ifn, jf) = 0.0;
                    }
                }
            }
            if (!(finemass >= dry_tolerance)) continue;
            for (int iv = 1; iv < 7; iv++) {       /* Fortran ivar = iv + 1 */
                int flag = 0;
                s1p = VC(iv, i + 1, j) - VC(iv, i, j);
                s1m = VC(iv, i, j) - VC(iv, i - 1, j);
                slopex = dmin(fabs(s1p), fabs(s1m)) * dsign(1.0, VC(iv, i + 1, j) - VC(iv, i - 1, j));
                if (s1m * s1p <= 0.0) slopex = 0.0;
                s1p = VC(iv, i, j + 1) - VC(iv, i, j);
                s1m = VC(iv, i, j) - VC(iv, i, j - 1);
                slopey = dmin(fabs(s1p), fabs(s1m)) * dsign(1.0, VC(iv, i, j + 1) - VC(iv, i, j - 1));
                if (s1m * s1p <= 0.0) slopey = 0.0;
                double velmax, velmin;
                if (VC(0, i, j) > dry_tolerance) {
                    velmax = VC(iv, i, j) / VC(0, i, j);
                    velmin = VC(iv, i, j) / VC(0, i, j);
                } else {
                    velmax = 0.0;
                    velmin = 0.0;
                    if (iv + 1 == 4) {
                        velmax = 0.0;
                        velmin = 0.0;
                    } else if (iv + 1 == 5) {
                        velmax = rho_f * grav;
                        velmin = rho_f * grav;
                    }
                }
                for (int ii = -1; ii <= 1; ii += 2) {
                    if (VC(0, i + ii, j) > dry_tolerance) {
                        const double vel = VC(iv, i + ii, j) / VC(0, i + ii, j);
                        velmax = dmax(vel, velmax);
                        velmin = dmin(vel, velmin);
                    }
                    if (VC(0, i, j + ii) > dry_tolerance) {
                        const double vel = VC(iv, i, j + ii) / VC(0, i, j + ii);
                        velmax = dmax(vel, velmax);
                        velmin = dmin(vel, velmin);
                    }
                }
                for (int ico = 1; ico <= rx; ico++) {      /* ico outer, jco inner (:330-331) */
                    if (fineflag1 || flag) break;
                    for (int jco = 1; jco <= ry; jco++) {
                        const int jf = (j - 1) * ry + nghost + jco - 1;
                        const int ifn = (i - 1) * rx + nghost + ico - 1;
                        const double yoff = ((double)jco - 0.5) / (double)ry - 0.5;
                        const double xoff = ((double)ico - 0.5) / (double)rx - 0.5;
                        const double hvf = (VC(iv, i, j) + xoff * slopex) + yoff * slopey;
                        const double vf = hvf / VF(0, ifn, jf);
                        if (vf > velmax || vf < velmin) {
                            flag = 1;
                            break;
                        }
                        VF(iv, ifn, jf) = hvf;
                    }
                }
                if (fineflag1 || flag) {
                    const double area = (double)(rx * ry);
                    const double dividemass = dmax(finemass, VC(0, i, j));
                    double vnew = area * VC(iv, i, j) / dividemass;
                    if (iv + 1 > 3) {
                        vnew = dmax(velmin, vnew);
                        vnew = dmin(velmax, vnew);
                    }
                    for (int jco = 1; jco <= ry; jco++)
                        for (int ico = 1; ico <= rx; ico++) {
                            const int jf = (j - 1) * ry + nghost + jco - 1;
                            const int ifn = (i - 1) * rx + nghost + ico - 1;
                            VF(iv, ifn, jf) = vnew * VF(0, ifn, jf);
                        }
                }
            }
        }
    }
#undef VC
#undef BCX
#undef VF
}

/* ------------------------------------------------------------------ filpatch
 * filpatch.f90:137-243 (the coarse rectangle around the unset cells),
 * Fortran integer division (truncation toward zero) as the reference has it. */
void oracle_amr_coarse_box(int isl, int isr, int jsb, int jst, int rx, int ry, int* box)
{
    const int nghost = 2;
    box[0] = (isl - rx + nghost * rx) / rx - nghost;   /* iplo */
    box[1] = (isr + rx) / rx;                          /* iphi */
    box[2] = (jsb - ry + nghost * ry) / ry - nghost;   /* jplo */
    box[3] = (jst + ry) / ry;                          /* jphi */
}

/* filpatch.f90:256-520: interpolation of the unset cells (flaguse == 0) of
 * the region ilo..ihi x jlo..jhi (level index space) from the coarse
 * rectangle valcrse (mxc x myc cells from (iplo, jplo)), force_dry and
 * variable_eta_init off (sea level constant), coordinate_system 1.  The
 * region's cell (ilo + a, jlo + b) sits at column r0 + a, row c0 + b of
 * valbig / aux (0-based, mx columns); flaguse is (ihi-ilo+1) x (jhi-jlo+1),
 * column fastest.  Loops as the reference has them: the depth over every
 * fine cell (j outer), then per component the coarse slopes and velocity
 * bounds, the fine cells, and the reset of flagged coarse cells.  Returns 0,
 * or -1 if an unset cell falls outside the interior of the coarse
 * rectangle. */
int oracle_amr_filpatch(const double* valcrse, const double* auxcrse, int mauxc, int mxc, int myc,
                        int iplo, int jplo, double* valbig, const double* aux, int maux, int mx,
                        int r0, int c0, int ilo, int ihi, int jlo, int jhi,
                        const unsigned char* flaguse, int rx, int ry, double xlower,
                        double ylower, double dx_fine, double dy_fine, double dx_coarse,
                        double dy_coarse, double sea_level, double dry_tolerance, double rho_f,
                        double grav)
{
    const int mx_patch = ihi - ilo + 1, my_patch = jhi - jlo + 1;
    const size_t nc = (size_t)mxc * (size_t)myc;
#define VCR(n, i, j) valcrse[(n) + 7 * ((size_t)((i) - 1) + (size_t)mxc * (size_t)((j) - 1))]
#define BCR(i, j) auxcrse[(size_t)mauxc * ((size_t)((i) - 1) + (size_t)mxc * (size_t)((j) - 1))]
#define VB(n, i, j) valbig[(n) + 7 * ((size_t)(r0 + (i) - 1) + (size_t)mx * (size_t)(c0 + (j) - 1))]
#define BF(i, j) aux[(size_t)maux * ((size_t)(r0 + (i) - 1) + (size_t)mx * (size_t)(c0 + (j) - 1))]
#define FU(i, j) flaguse[(size_t)((i) - 1) + (size_t)mx_patch * (size_t)((j) - 1)]
#define CX(a, i, j) a[(size_t)((i) - 1) + (size_t)mxc * (size_t)((j) - 1)]
    double* eta_coarse = calloc(nc, sizeof(double));
    double* slope1 = calloc(nc, sizeof(double));
    double* slope2 = calloc(nc, sizeof(double));
    double* fine_mass = calloc(nc, sizeof(double));
    double* vel_max = calloc(nc, sizeof(double));
    double* vel_min = calloc(nc, sizeof(double));
    int* fine_cell_count = calloc(nc, sizeof(int));
    unsigned char* flag1 = calloc(nc, 1);
    unsigned char* flagn = calloc(nc, 1);
    int rc = 0;
    const double xlow_coarse = xlower + iplo * dx_coarse;
    const double ylow_coarse = ylower + jplo * dy_coarse;

    /* :274-287 surface with dry limiting */
    for (int j = 1; j <= myc; j++)
        for (int i = 1; i <= mxc; i++) {
            const double h = VCR(0, i, j), b = BCR(i, j);
            if (h < dry_tolerance) CX(eta_coarse, i, j) = sea_level;
            else CX(eta_coarse, i, j) = h + b;
        }
    /* :294-313 limited gradients of eta */
    for (int j = 2; j <= myc - 1; j++)
        for (int i = 2; i <= mxc - 1; i++) {
            double down = CX(eta_coarse, i, j) - CX(eta_coarse, i - 1, j);
            double up = CX(eta_coarse, i + 1, j) - CX(eta_coarse, i, j);
            if (up * down > 0.0)
                CX(slope1, i, j) = dmin(fabs(up), fabs(down)) *
                                   dsign(1.0, CX(eta_coarse, i + 1, j) - CX(eta_coarse, i - 1, j));
            down = CX(eta_coarse, i, j) - CX(eta_coarse, i, j - 1);
            up = CX(eta_coarse, i, j + 1) - CX(eta_coarse, i, j);
            if (up * down > 0.0)
                CX(slope2, i, j) = dmin(fabs(up), fabs(down)) *
                                   dsign(1.0, CX(eta_coarse, i, j + 1) - CX(eta_coarse, i, j - 1));
        }
    const double ratio_y = (double)ry, ratio_x = (double)rx;
    int reloop = 0;
    /* :317-376 depth */
    for (int jf = 1; jf <= my_patch; jf++) {
        const int jc = (int)floor((double)(jf + jlo - 1) / ratio_y) - jplo + 1;
        const double ycent_coarse = ylow_coarse + ((double)jc - 0.5) * dy_coarse;
        const double ycent_fine = ylower + ((double)(jf - 1 + jlo) + 0.5) * dy_fine;
        const double eta2 = (ycent_fine - ycent_coarse) / dy_coarse;
        for (int ifn = 1; ifn <= mx_patch; ifn++) {
            const int ic = (int)floor((double)(ifn + ilo - 1) / ratio_x) - iplo + 1;
            const double xcent_coarse = xlow_coarse + ((double)ic - 0.5) * dx_coarse;
            const double xcent_fine = xlower + ((double)(ifn - 1 + ilo) + 0.5) * dx_fine;
            const double eta1 = (xcent_fine - xcent_coarse) / dx_coarse;
            if (FU(ifn, jf) != 0) continue;
            if (ic < 2 || ic > mxc - 1 || jc < 2 || jc > myc - 1) { rc = -1; goto done; }
            CX(fine_cell_count, ic, jc) += 1;
            const double eta_fine = (CX(eta_coarse, ic, jc) + eta1 * CX(slope1, ic, jc)) +
                                    eta2 * CX(slope2, ic, jc);
            const double h_fine = dmax(eta_fine - BF(ifn, jf), 0.0);
            VB(0, ifn, jf) = h_fine;
            CX(fine_mass, ic, jc) = CX(fine_mass, ic, jc) + h_fine;
            if (h_fine < dry_tolerance) {
                CX(flag1, ic, jc) = 1;
                reloop = 1;
            }
        }
    }
    /* :379-516 the other components */
    for (int n = 2; n <= 7; n++) {
        const int m = n - 1;
        memset(slope1, 0, nc * sizeof(double));
        memset(slope2, 0, nc * sizeof(double));
        memset(flagn, 0, nc);
        for (int j = 2; j <= myc - 1; j++)
            for (int i = 2; i <= mxc - 1; i++) {
                double down = VCR(m, i, j) - VCR(m, i - 1, j);
                double up = VCR(m, i + 1, j) - VCR(m, i, j);
                if (up * down > 0.0)
                    CX(slope1, i, j) = dmin(fabs(up), fabs(down)) *
                                       dsign(1.0, VCR(m, i + 1, j) - VCR(m, i - 1, j));
                down = VCR(m, i, j) - VCR(m, i, j - 1);
                up = VCR(m, i, j + 1) - VCR(m, i, j);
                if (up * down > 0.0)
                    CX(slope2, i, j) = dmin(fabs(up), fabs(down)) *
                                       dsign(1.0, VCR(m, i, j + 1) - VCR(m, i, j - 1));
                if (VCR(0, i, j) > dry_tolerance) {
                    CX(vel_max, i, j) = VCR(m, i, j) / VCR(0, i, j);
                    CX(vel_min, i, j) = VCR(m, i, j) / VCR(0, i, j);
                } else {
                    CX(vel_min, i, j) = 0.0;
                    CX(vel_max, i, j) = 0.0;
                    if (n == 4) {
                        CX(vel_max, i, j) = 0.0;
                        CX(vel_min, i, j) = 0.0;
                    } else if (n == 5) {
                        CX(vel_max, i, j) = rho_f * grav;
                        CX(vel_min, i, j) = rho_f * grav;
                    }
                }
                for (int s = -1; s <= 1; s += 2) {
                    if (VCR(0, i + s, j) > dry_tolerance) {
                        CX(vel_max, i, j) = dmax(CX(vel_max, i, j), VCR(m, i + s, j) / VCR(0, i + s, j));
                        CX(vel_min, i, j) = dmin(CX(vel_min, i, j), VCR(m, i + s, j) / VCR(0, i + s, j));
                    }
                    if (VCR(0, i, j + s) > dry_tolerance) {
                        CX(vel_max, i, j) = dmax(CX(vel_max, i, j), VCR(m, i, j + s) / VCR(0, i, j + s));
                        CX(vel_min, i, j) = dmin(CX(vel_min, i, j), VCR(m, i, j + s) / VCR(0, i, j + s));
                    }
                }
            }
        for (int jf = 1; jf <= my_patch; jf++) {
            const int jc = (int)floor((double)(jf + jlo - 1) / ratio_y) - jplo + 1;
            const double ycent_coarse = ylow_coarse + ((double)jc - 0.5) * dy_coarse;
            const double ycent_fine = ylower + ((double)(jf - 1 + jlo) + 0.5) * dy_fine;
            const double eta2 = (ycent_fine - ycent_coarse) / dy_coarse;
            for (int ifn = 1; ifn <= mx_patch; ifn++) {
                const int ic = (int)floor((double)(ifn + ilo - 1) / ratio_x) - iplo + 1;
                const double xcent_coarse = xlow_coarse + ((double)ic - 0.5) * dx_coarse;
                const double xcent_fine = xlower + ((double)(ifn - 1 + ilo) + 0.5) * dx_fine;
                const double eta1 = (xcent_fine - xcent_coarse) / dx_coarse;
                if (FU(ifn, jf) != 0 || CX(flag1, ic, jc)) continue;
                const double hv_fine = (VCR(m, ic, jc) + eta1 * CX(slope1, ic, jc)) +
                                       eta2 * CX(slope2, ic, jc);
                const double v_fine = hv_fine / VB(0, ifn, jf);
                if (v_fine < CX(vel_min, ic, jc) || v_fine > CX(vel_max, ic, jc)) {
                    CX(flagn, ic, jc) = 1;
                    reloop = 1;
                } else {
                    VB(m, ifn, jf) = hv_fine;
                }
            }
        }
        if (reloop) {
            for (int jf = 1; jf <= my_patch; jf++) {
                const int jc = (int)floor((double)(jf + jlo - 1) / ratio_y) - jplo + 1;
                for (int ifn = 1; ifn <= mx_patch; ifn++) {
                    const int ic = (int)floor((double)(ifn + ilo - 1) / ratio_x) - iplo + 1;
                    if (FU(ifn, jf) != 0) continue;
                    if (!(CX(flag1, ic, jc) || CX(flagn, ic, jc))) continue;
                    if (CX(fine_mass, ic, jc) > dry_tolerance) {
                        const double h_coarse = VCR(0, ic, jc);
                        const double h_count = (double)CX(fine_cell_count, ic, jc);
                        const double h_fine_average = CX(fine_mass, ic, jc) / h_count;
                        const double divide_mass = dmax(h_coarse, h_fine_average);
                        double v_new = VCR(m, ic, jc) / divide_mass;
                        if (n > 3) {
                            v_new = dmax(CX(vel_min, ic, jc), v_new);
                            v_new = dmin(CX(vel_max, ic, jc), v_new);
                        }
                        VB(m, ifn, jf) = v_new * VB(0, ifn, jf);
                    } else {
                        VB(m, ifn, jf) = 0.0;
                    }
                }
            }
        }
    }
done:
    free(eta_coarse); free(slope1); free(slope2); free(fine_mass); free(vel_max); free(vel_min);
    free(fine_cell_count); free(flag1); free(flagn);
#undef VCR
#undef BCR
#undef VB
#undef BF
#undef FU
#undef CX
    return rc;
}

/* ------------------------------------------------ region fill (AMRClaw steps)
 * Not in the reference: AMRClaw 5.x intfil/icall (copy from the grids of a
 * level, linear in time between a grid's two stored times), trimbd
 * (bounding box of the unset cells) and bc2amr (extrap / wall fill of the
 * cells outside the domain), restated from their published behaviour.
 * Region layout: cell (ilo + a, jlo + b) at column r0 + a, row c0 + b of
 * q(7, mx, .) and aux(maux, mx, .); flag (ihi-ilo+1) x (jhi-jlo+1). */
void oracle_amr_intfil(double* q, double* aux, int maux, int mx, int r0, int c0, int ilo,
                       int ihi, int jlo, int jhi, unsigned char* flag, const double* sq,
                       const double* sq_old, double w_old, const double* saux, int smaux,
                       int silo, int sjlo, int smx, int smy, int copy_aux)
{
    const int mxp = ihi - ilo + 1;
    for (int j = jlo; j <= jhi; j++)
        for (int i = ilo; i <= ihi; i++) {
            if (i < silo || i > silo + smx - 1 || j < sjlo || j > sjlo + smy - 1) continue;
            flag[(size_t)(i - ilo) + (size_t)mxp * (j - jlo)] = 1;
            const size_t d = (size_t)(r0 + i - ilo) + (size_t)mx * (size_t)(c0 + j - jlo);
            const size_t s = (size_t)(i - silo + 2) + (size_t)(smx + 4) * (size_t)(j - sjlo + 2);
            for (int m = 0; m < 7; m++) {
                if (sq_old && w_old != 0.0)
                    q[7 * d + m] = sq[7 * s + m] * (1.0 - w_old) + sq_old[7 * s + m] * w_old;
                else
                    q[7 * d + m] = sq[7 * s + m];
            }
            if (copy_aux)
                for (int k = 0; k < (maux < smaux ? maux : smaux); k++)
                    aux[(size_t)maux * d + k] = saux[(size_t)smaux * s + k];
        }
}

/* trimbd: box = (isl, isr, jsb, jst) of the unset cells in level index
 * space; returns 1 if every cell is set (box untouched), else 0. */
int oracle_amr_trimbd(const unsigned char* flag, int ilo, int ihi, int jlo, int jhi, int* box)
{
    const int mxp = ihi - ilo + 1;
    int isl = ihi + 1, isr = ilo - 1, jsb = jhi + 1, jst = jlo - 1;
    for (int j = jlo; j <= jhi; j++)
        for (int i = ilo; i <= ihi; i++)
            if (flag[(size_t)(i - ilo) + (size_t)mxp * (j - jlo)] == 0) {
                if (i < isl) isl = i;
                if (i > isr) isr = i;
                if (j < jsb) jsb = j;
                if (j > jst) jst = j;
            }
    if (isl > isr) return 1;
    box[0] = isl; box[1] = isr; box[2] = jsb; box[3] = jst;
    return 0;
}

/* bc2amr on a region: first the columns outside the domain in x (rows
 * inside in y) from the x rule, then the rows outside in y (every column)
 * from the y rule, as the patch ghost fill orders them; q and every aux
 * component copied, the normal momentum negated at walls (bc 3). */
void oracle_amr_bc2amr(double* q, double* aux, int maux, int mx, int r0, int c0, int ilo,
                       int ihi, int jlo, int jhi, int iregsz, int jregsz, const int* bc)
{
#define RC(i, j) ((size_t)(r0 + (i) - ilo) + (size_t)mx * (size_t)(c0 + (j) - jlo))
    for (int j = jlo; j <= jhi; j++) {
        if (j < 0 || j >= jregsz) continue;
        for (int i = ilo; i <= ihi; i++) {
            int si;
            double f = 1.0;
            if (i < 0) {
                if (bc[0] == 3) { si = -1 - i; f = -1.0; } else si = 0;
            } else if (i >= iregsz) {
                if (bc[1] == 3) { si = 2 * iregsz - 1 - i; f = -1.0; } else si = iregsz - 1;
            } else {
                continue;
            }
            for (int m = 0; m < 7; m++)
                q[7 * RC(i, j) + m] = (m == 1 ? f : 1.0) * q[7 * RC(si, j) + m];
            for (int k = 0; k < maux; k++) aux[(size_t)maux * RC(i, j) + k] = aux[(size_t)maux * RC(si, j) + k];
        }
    }
    for (int j = jlo; j <= jhi; j++) {
        int sj;
        double f = 1.0;
        if (j < 0) {
            if (bc[2] == 3) { sj = -1 - j; f = -1.0; } else sj = 0;
        } else if (j >= jregsz) {
            if (bc[3] == 3) { sj = 2 * jregsz - 1 - j; f = -1.0; } else sj = jregsz - 1;
        } else {
            continue;
        }
        for (int i = ilo; i <= ihi; i++) {
            for (int m = 0; m < 7; m++)
                q[7 * RC(i, j) + m] = (m == 2 ? f : 1.0) * q[7 * RC(i, sj) + m];
            for (int k = 0; k < maux; k++) aux[(size_t)maux * RC(i, j) + k] = aux[(size_t)maux * RC(i, sj) + k];
        }
    }
#undef RC
}

/* setaux.f90:59-258 for a patch (interior cells ilo.., jlo.., mx x my, 2
 * ghost cells), coordinate_system 1, no auxinit files: aux(1) from bed
 * where bed_set is 0 and the cell centre lies in the domain (the topography
 * loop, :75-139, with cellgridintegrate's result supplied), then the rows
 * and the columns whose centres lie outside the domain copy aux(1) from the
 * cell the ceiling shifts name (:154-184), aux(2..maux) = 0 (:190-198),
 * aux(i_phi = 2) = phi (:260-264), aux(i_theta = 3) = theta (:267-273). */
void oracle_amr_setaux(double* aux, int maux, int mx, int my, int ilo, int jlo,
                       const unsigned char* bed_set, const double* bed, double xlower,
                       double xupper, double ylower, double yupper, double dx, double dy,
                       double phi, double theta)
{
    const int mbc = 2, nx = mx + 4;
#define A1(ii, jj) aux[(size_t)maux * ((size_t)((ii) + mbc - 1) + (size_t)nx * (size_t)((jj) + mbc - 1))]
#define CI(ii, jj) ((size_t)((ii) + mbc - 1) + (size_t)nx * (size_t)((jj) + mbc - 1))
    /* ilo, jlo here are the level indices of the first interior cell, which
     * is setaux's floor((xlow - xlower + 0.05 dx)/dx) */
    for (int jj = 1 - mbc; jj <= my + mbc; jj++) {
        const double ym = ylower + (jlo + jj - 1.0) * dy, yp = ylower + (jlo + jj) * dy;
        const double y = 0.5 * (ym + yp);
        for (int ii = 1 - mbc; ii <= mx + mbc; ii++) {
            const double xm = xlower + (ilo + ii - 1.0) * dx, xp = xlower + (ilo + ii) * dx;
            const double x = 0.5 * (xm + xp);
            if (bed_set[CI(ii, jj)]) continue;
            if (y > yupper || y < ylower || x > xupper || x < xlower) continue;
            A1(ii, jj) = bed[CI(ii, jj)];
        }
    }
    for (int jj = 1 - mbc; jj <= my + mbc; jj++) {
        const double y = ylower + (jlo + jj - 0.5) * dy;
        if (y < ylower || y > yupper) {
            for (int ii = 1 - mbc; ii <= mx + mbc; ii++) {
                const double x = xlower + (ilo + ii - 0.5) * dx;
                const int iint = ii + (int)dmax(0.0, ceil((xlower - x) / dx)) -
                                 (int)dmax(0.0, ceil((x - xupper) / dx));
                const int jint = jj + (int)dmax(0.0, ceil((ylower - y) / dy)) -
                                 (int)dmax(0.0, ceil((y - yupper) / dy));
                A1(ii, jj) = A1(iint, jint);
            }
        }
    }
    for (int ii = 1 - mbc; ii <= mx + mbc; ii++) {
        const double x = xlower + (ilo + ii - 0.5) * dx;
        if (x < xlower || x > xupper) {
            for (int jj = 1 - mbc; jj <= my + mbc; jj++) {
                const double y = ylower + (jlo + jj - 0.5) * dy;
                const int iint = ii + (int)dmax(0.0, ceil((xlower - x) / dx)) -
                                 (int)dmax(0.0, ceil((x - xupper) / dx));
                const int jint = jj + (int)dmax(0.0, ceil((ylower - y) / dy)) -
                                 (int)dmax(0.0, ceil((y - yupper) / dy));
                A1(ii, jj) = A1(iint, jint);
            }
        }
    }
    for (int jj = 1 - mbc; jj <= my + mbc; jj++)
        for (int ii = 1 - mbc; ii <= mx + mbc; ii++) {
            double* a = aux + (size_t)maux * CI(ii, jj);
            for (int k = 1; k < maux; k++) a[k] = 0.0;
            a[1] = phi;
            a[2] = theta;
        }
#undef A1
#undef CI
}
