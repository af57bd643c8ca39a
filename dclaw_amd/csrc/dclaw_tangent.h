// dclaw_tangent.h -- the tangent-linear (forward-mode) step: the strip body
// instantiated over a value with one tangent.
//
// Every operation computes its value exactly as the double step does (same
// operations, same order, same elementary functions), so the value part of a
// tangent-linear step is the step itself, bit for bit; the tangent follows
// the chain rule.  Where the step branches (comparisons, max/min, |x|, the
// dry tests) the tangent follows the branch the value takes: the derivative
// of the branch that is active.  The tangent of dt is not carried (a fixed
// dt sequence, the one the values choose).
//
// Inputs: the state q with its tangent dq (the perturbation of the initial
// state), and the parameter record with a tangent record dP (the derivative
// of each parameter along the direction of differentiation: 1 in phi alone
// gives d/dphi).  Output: q_{n+1} and dq_{n+1} = J_q dq_n + J_P dP.
#pragma once

#include "dclaw_strip.h"

namespace dclaw_tl {

struct Dual {
    double v, d;
    Dual() = default;
    DC_HD Dual(double x) : v(x), d(0.0) {}
    DC_HD Dual(double x, double dx) : v(x), d(dx) {}
};

DC_HD Dual operator+(Dual a) { return a; }
DC_HD Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
DC_HD Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
DC_HD Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
DC_HD Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
DC_HD Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
DC_HD Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
DC_HD Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
DC_HD Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
DC_HD Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
DC_HD Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
DC_HD Dual operator/(Dual a, Dual b)
{
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
DC_HD Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
DC_HD Dual operator/(double a, Dual b)
{
    const double q = a / b.v;
    return Dual(q, -(q * b.d) / b.v);
}
DC_HD Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
DC_HD Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
DC_HD Dual& operator*=(Dual& a, Dual b) { return a = a * b; }
DC_HD Dual& operator/=(Dual& a, Dual b) { return a = a / b; }

#define DCLAW_TL_CMP(op)                                                     \
    DC_HD bool operator op(Dual a, Dual b) { return a.v op b.v; }            \
    DC_HD bool operator op(Dual a, double b) { return a.v op b; }            \
    DC_HD bool operator op(double a, Dual b) { return a op b.v; }
DCLAW_TL_CMP(<)
DCLAW_TL_CMP(>)
DCLAW_TL_CMP(<=)
DCLAW_TL_CMP(>=)
DCLAW_TL_CMP(==)
DCLAW_TL_CMP(!=)
#undef DCLAW_TL_CMP

DC_HD Dual sqrt(Dual a)
{
    const double s = ::sqrt(a.v);
    return Dual(s, s > 0.0 ? a.d / (2.0 * s) : 0.0);
}
DC_HD Dual fabs(Dual a) { return Dual(::fabs(a.v), signbit(a.v) ? -a.d : a.d); }
DC_HD Dual copysign(Dual a, Dual b)
{
    const double c = ::copysign(a.v, b.v);
    return Dual(c, signbit(c) != signbit(a.v) ? -a.d : a.d);
}
DC_HD Dual dcm_exp(Dual a)
{
    const double e = ::dcm_exp(a.v);
    return Dual(e, e * a.d);
}
DC_HD Dual dcm_log(Dual a) { return Dual(::dcm_log(a.v), a.d / a.v); }
DC_HD Dual dcm_atan(Dual a) { return Dual(::dcm_atan(a.v), a.d / (1.0 + a.v * a.v)); }
DC_HD Dual dcm_tan(Dual a)
{
    const double t = ::dcm_tan(a.v);
    return Dual(t, a.d * (1.0 + t * t));
}
DC_HD Dual dcm_tanh(Dual a)
{
    const double t = ::dcm_tanh(a.v);
    return Dual(t, a.d * (1.0 - t * t));
}
DC_HD Dual dcm_pow(Dual a, Dual b)
{
    const double p = ::dcm_pow(a.v, b.v);
    double d = 0.0;
    if (a.d != 0.0) d = d + b.v * ::dcm_pow(a.v, b.v - 1.0) * a.d;
    if (b.d != 0.0 && a.v > 0.0) d = d + p * ::dcm_log(a.v) * b.d;
    return Dual(p, d);
}

// The parameter record with every floating field carrying its tangent.
struct Params {
    Dual grav, dry_tolerance, speed_limit;
    Dual rho_s, rho_f, m_crit, mref, kref;
    Dual phi;
    Dual delta, mu, alpha_c, c1, sigma_0;
    Dual kappa;
    Dual beta_seg, delta_kr_order;
    Dual manning_coefficient;
    Dual friction_depth, manning_max;
    int src2method;
    int alphamethod, rarecorrectortest, segregation;
    int friction_forcing, dd_manning;
    int order;
    int transverse;
    int mthlim[5];
    int mwaves;
    Dual entrainment_rate, me, chie;
    int entrainment;
    int entrainment_method;
    int curvature;
};

DC_HD Params make_params(const dclaw_params& p, const dclaw_params& dp)
{
    Params o;
#define F(x) o.x = Dual(p.x, dp.x)
    F(grav); F(dry_tolerance); F(speed_limit);
    F(rho_s); F(rho_f); F(m_crit); F(mref); F(kref);
    F(phi);
    F(delta); F(mu); F(alpha_c); F(c1); F(sigma_0);
    F(kappa);
    F(beta_seg); F(delta_kr_order);
    F(manning_coefficient);
    F(friction_depth); F(manning_max);
    F(entrainment_rate); F(me); F(chie);
#undef F
    o.src2method = p.src2method;
    o.alphamethod = p.alphamethod;
    o.rarecorrectortest = p.rarecorrectortest;
    o.segregation = p.segregation;
    o.friction_forcing = p.friction_forcing;
    o.dd_manning = p.dd_manning;
    o.order = p.order;
    o.transverse = p.transverse;
    for (int k = 0; k < 5; k++) o.mthlim[k] = p.mthlim[k];
    o.mwaves = p.mwaves;
    o.entrainment = p.entrainment;
    o.entrainment_method = p.entrainment_method;
    o.curvature = p.curvature;
    return o;
}

using Real = Dual;
// src2method 2 is not differentiated (src2_cell writes NaN for it here)
constexpr bool kFe4quad = false;

#include "dclaw_physics_body.inc"

using dclaw::aidx;
using dclaw::bc_source;
using dclaw::CLASSIFY_CHUNK;
using dclaw::DRY_CHUNK;
using dclaw::MODE_FLUX;
using dclaw::MODE_STEP;
using dclaw::MODE_STEP_NOSRC2;
using dclaw::qidx;
using dclaw::STRIP_HALO;
using dclaw::STRIP_LANES;
using dclaw::STRIP_W;

// A q array with its tangent array (same layout).
struct DIn {
    const double* v;
    const double* d;
};
struct DOut {
    double* v;
    double* d;
};

// Arguments of the tangent-linear strip body: as dclaw::StripArgs, with q in
// and out carrying tangents.  aux (the bed and, with entrainment, h_e) has no
// tangent; aux_out and the step2 fluxes are not produced by this step.
struct StripArgs {
    DIn q_in;
    DOut q_out;
    const double* aux;
    double* aux_out;
    int maux, mx, my;
    double dx, dy, dt;
    int i0, jA, jB;
    int bc[4];
    double *fm, *fp, *gm, *gp;
    int* tile_flag;
    bool known_wet = false;
};

DC_HD Dual ld(const DIn& p, size_t i) { return Dual(p.v[i], p.d[i]); }
DC_HD void st(const DOut& p, size_t i, Dual x)
{
    p.v[i] = x.v;
    p.d[i] = x.d;
}
DC_HD void st(double* p, size_t i, Dual x) { p[i] = x.v; }

#include "dclaw_strip_body.inc"

}  // namespace dclaw_tl
