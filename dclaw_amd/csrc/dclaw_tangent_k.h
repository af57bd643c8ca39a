// dclaw_tangent_k.h -- the point physics over a value with K tangents at
// once: forward mode in K directions per evaluation.
//
// Each tangent component follows exactly the arithmetic of the one-tangent
// type dclaw_tl::Dual (same operations, same order, same conventions at
// kinks: the branch the value takes, sqrt'(0) = 0, the sign rules of |x| and
// copysign, pow's two guarded terms), so component k of one evaluation is,
// bit for bit, the tangent of a dclaw_tl evaluation seeded with direction k.
// The value and every elementary function's value (exp, atan, tan, tanh,
// pow, sqrt, division) are computed once for all K directions.  The adjoint
// step uses it for the stages whose vector-Jacobian products are seeded
// (prepare, calc_taudir, src2).
#pragma once

#include "dclaw_tangent.h"

namespace dclaw_tlk {

constexpr int K = 4;  // directions per evaluation

struct DualK {
    double v, d[K];
    DualK() = default;
    DC_HD DualK(double x) : v(x)
    {
#pragma unroll
        for (int k = 0; k < K; k++) d[k] = 0.0;
    }
};

#define DCLAW_TLK_MAP(expr)                                                  \
    _Pragma("unroll") for (int k = 0; k < K; k++) o.d[k] = (expr);

DC_HD DualK operator+(DualK a) { return a; }
DC_HD DualK operator-(DualK a)
{
    DualK o;
    o.v = -a.v;
    DCLAW_TLK_MAP(-a.d[k])
    return o;
}
DC_HD DualK operator+(DualK a, DualK b)
{
    DualK o;
    o.v = a.v + b.v;
    DCLAW_TLK_MAP(a.d[k] + b.d[k])
    return o;
}
DC_HD DualK operator+(DualK a, double b)
{
    DualK o = a;
    o.v = a.v + b;
    return o;
}
DC_HD DualK operator+(double a, DualK b)
{
    DualK o = b;
    o.v = a + b.v;
    return o;
}
DC_HD DualK operator-(DualK a, DualK b)
{
    DualK o;
    o.v = a.v - b.v;
    DCLAW_TLK_MAP(a.d[k] - b.d[k])
    return o;
}
DC_HD DualK operator-(DualK a, double b)
{
    DualK o = a;
    o.v = a.v - b;
    return o;
}
DC_HD DualK operator-(double a, DualK b)
{
    DualK o;
    o.v = a - b.v;
    DCLAW_TLK_MAP(-b.d[k])
    return o;
}
DC_HD DualK operator*(DualK a, DualK b)
{
    DualK o;
    o.v = a.v * b.v;
    DCLAW_TLK_MAP(a.d[k] * b.v + a.v * b.d[k])
    return o;
}
DC_HD DualK operator*(DualK a, double b)
{
    DualK o;
    o.v = a.v * b;
    DCLAW_TLK_MAP(a.d[k] * b)
    return o;
}
DC_HD DualK operator*(double a, DualK b)
{
    DualK o;
    o.v = a * b.v;
    DCLAW_TLK_MAP(a * b.d[k])
    return o;
}
DC_HD DualK operator/(DualK a, DualK b)
{
    DualK o;
    const double q = a.v / b.v;
    o.v = q;
    DCLAW_TLK_MAP((a.d[k] - q * b.d[k]) / b.v)
    return o;
}
DC_HD DualK operator/(DualK a, double b)
{
    DualK o;
    o.v = a.v / b;
    DCLAW_TLK_MAP(a.d[k] / b)
    return o;
}
DC_HD DualK operator/(double a, DualK b)
{
    DualK o;
    const double q = a / b.v;
    o.v = q;
    DCLAW_TLK_MAP(-(q * b.d[k]) / b.v)
    return o;
}
DC_HD DualK& operator+=(DualK& a, DualK b) { return a = a + b; }
DC_HD DualK& operator-=(DualK& a, DualK b) { return a = a - b; }
DC_HD DualK& operator*=(DualK& a, DualK b) { return a = a * b; }
DC_HD DualK& operator/=(DualK& a, DualK b) { return a = a / b; }

#define DCLAW_TLK_CMP(op)                                                     \
    DC_HD bool operator op(DualK a, DualK b) { return a.v op b.v; }           \
    DC_HD bool operator op(DualK a, double b) { return a.v op b; }            \
    DC_HD bool operator op(double a, DualK b) { return a op b.v; }
DCLAW_TLK_CMP(<)
DCLAW_TLK_CMP(>)
DCLAW_TLK_CMP(<=)
DCLAW_TLK_CMP(>=)
DCLAW_TLK_CMP(==)
DCLAW_TLK_CMP(!=)
#undef DCLAW_TLK_CMP

DC_HD DualK sqrt(DualK a)
{
    DualK o;
    const double s = ::sqrt(a.v);
    o.v = s;
    DCLAW_TLK_MAP(s > 0.0 ? a.d[k] / (2.0 * s) : 0.0)
    return o;
}
DC_HD DualK fabs(DualK a)
{
    DualK o;
    o.v = ::fabs(a.v);
    const bool neg = signbit(a.v);
    DCLAW_TLK_MAP(neg ? -a.d[k] : a.d[k])
    return o;
}
DC_HD DualK copysign(DualK a, DualK b)
{
    DualK o;
    const double c = ::copysign(a.v, b.v);
    o.v = c;
    const bool flip = signbit(c) != signbit(a.v);
    DCLAW_TLK_MAP(flip ? -a.d[k] : a.d[k])
    return o;
}
DC_HD DualK dcm_exp(DualK a)
{
    DualK o;
    const double e = ::dcm_exp(a.v);
    o.v = e;
    DCLAW_TLK_MAP(e * a.d[k])
    return o;
}
DC_HD DualK dcm_log(DualK a)
{
    DualK o;
    o.v = ::dcm_log(a.v);
    DCLAW_TLK_MAP(a.d[k] / a.v)
    return o;
}
DC_HD DualK dcm_atan(DualK a)
{
    DualK o;
    o.v = ::dcm_atan(a.v);
    DCLAW_TLK_MAP(a.d[k] / (1.0 + a.v * a.v))
    return o;
}
DC_HD DualK dcm_tan(DualK a)
{
    DualK o;
    const double t = ::dcm_tan(a.v);
    o.v = t;
    DCLAW_TLK_MAP(a.d[k] * (1.0 + t * t))
    return o;
}
DC_HD DualK dcm_tanh(DualK a)
{
    DualK o;
    const double t = ::dcm_tanh(a.v);
    o.v = t;
    DCLAW_TLK_MAP(a.d[k] * (1.0 - t * t))
    return o;
}
DC_HD DualK dcm_pow(DualK a, DualK b)
{
    DualK o;
    const double p = ::dcm_pow(a.v, b.v);
    o.v = p;
    bool any_a = false, any_b = false;
    for (int k = 0; k < K; k++) {
        any_a = any_a || a.d[k] != 0.0;
        any_b = any_b || b.d[k] != 0.0;
    }
    const double pa = any_a ? ::dcm_pow(a.v, b.v - 1.0) : 0.0;
    const double lb = any_b && a.v > 0.0 ? ::dcm_log(a.v) : 0.0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double d = 0.0;
        if (a.d[k] != 0.0) d = d + b.v * pa * a.d[k];
        if (b.d[k] != 0.0 && a.v > 0.0) d = d + p * lb * b.d[k];
        o.d[k] = d;
    }
    return o;
}
#undef DCLAW_TLK_MAP

// The parameter record with every floating field carrying K tangents.
struct Params {
    DualK grav, dry_tolerance, speed_limit;
    DualK rho_s, rho_f, m_crit, mref, kref;
    DualK phi;
    DualK delta, mu, alpha_c, c1, sigma_0;
    DualK kappa;
    DualK beta_seg, delta_kr_order;
    DualK manning_coefficient;
    DualK friction_depth, manning_max;
    int src2method;
    int alphamethod, rarecorrectortest, segregation;
    int friction_forcing, dd_manning;
    int order;
    int transverse;
    int mthlim[5];
    int mwaves;
    DualK entrainment_rate, me, chie;
    int entrainment;
    int entrainment_method;
    int curvature;
};

// Values p; tangent k is dp[k]'s field, or 0 where dp[k] is null.
DC_HD Params make_params(const dclaw_params& p, const dclaw_params* const* dp)
{
    Params o;
#define F(x)                                                                 \
    o.x = DualK(p.x);                                                        \
    for (int k = 0; k < K; k++) o.x.d[k] = dp[k] ? dp[k]->x : 0.0;
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

using Real = DualK;
// src2method 2 is not differentiated (src2_cell writes NaN for it here)
constexpr bool kFe4quad = false;

#include "dclaw_physics_body.inc"
#include "dclaw_prepare_body.inc"

}  // namespace dclaw_tlk
