/*
 * dcmath.h -- portable, deterministic FP64 elementary functions.
 *
 * exp, expm1 (|x| <= 1), tanh, atan, tan, log and pow for positive bases,
 * built only from IEEE-754 double +, -, *, /, rint and exponent bit
 * manipulation, evaluated in a fixed order.  Compiled without FMA
 * contraction, every caller that includes this file gets bit-identical
 * results on the host CPU and on gfx950, which is what lets the GPU step
 * reproduce a CPU run of the same scheme bit for bit over thousands of
 * steps (the D-Claw scheme amplifies 1-ulp differences).
 *
 * Methods (textbook; constants are the mathematical values, split where
 * noted so the products with small integers are exact):
 *   exp   Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, with ln2 split
 *         in a 32-bit head and a tail; Taylor series of degree 13 in Horner
 *         form; scaling by 2^k through the exponent field (two steps into
 *         the subnormal range).
 *   atan  reduction to |t| < 7/16 at the break points 7/16, 11/16, 19/16,
 *         39/16 with atan(1/2), atan(1), atan(3/2), pi/2 split in head and
 *         tail; odd Taylor series to t^49.
 *   tan   reduction by k pi/2 with pi/2 split in three 33-bit pieces; Taylor
 *         series of sin and cos on |r| <= pi/4; tan = sin/cos, or -cos/sin
 *         for odd k.
 *   tanh  1 - 2/(exp(2|x|)+1) for |x| >= 1/2, e/(e+2) with e = expm1(2|x|)
 *         (Taylor, degree 20) below; 1 beyond |x| > 22.
 *   log   x = 2^e f, f in [sqrt(1/2), sqrt(2)); log f = 2 atanh(s),
 *         s = (f-1)/(f+1), odd series to s^23.
 *   pow   exp(y log x) for x > 0 (used only by optional terms).
 * Accuracy against a 64-bit-mantissa reference: a few ulp at most (tests).
 *
 * The includer defines DCM_FN (the function qualifiers) before including.
 */
#ifndef DCMATH_H
#define DCMATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

#ifndef DCM_FN
#error "define DCM_FN (function qualifiers) before including dcmath.h"
#endif

DCM_FN double dcm_from_bits(uint64_t b)
{
    double d;
    memcpy(&d, &b, sizeof d);
    return d;
}

DCM_FN uint64_t dcm_bits(double d)
{
    uint64_t b;
    memcpy(&b, &d, sizeof b);
    return b;
}

/* 2^k for k in [-1022, 1023] */
DCM_FN double dcm_pow2i(int k) { return dcm_from_bits((uint64_t)(k + 1023) << 52); }

#define DCM_LN2_HI 0x1.62e42fee00000p-1
#define DCM_LN2_LO 0x1.a39ef35793c76p-33
#define DCM_LOG2E 0x1.71547652b82fep+0

/* (e^r - 1)/r - 1 ... evaluated as e^r = 1 + r*(1 + r*(1/2 + r*(1/6 + ...))) */
DCM_FN double dcm_exp_poly(double r)
{
    double p = 0x1.6124613a86d09p-33;          /* 1/13! */
    p = 0x1.1eed8eff8d898p-29 + r * p;          /* 1/12! */
    p = 0x1.ae64567f544e4p-26 + r * p;          /* 1/11! */
    p = 0x1.27e4fb7789f5cp-22 + r * p;          /* 1/10! */
    p = 0x1.71de3a556c734p-19 + r * p;          /* 1/9!  */
    p = 0x1.a01a01a01a01ap-16 + r * p;          /* 1/8!  */
    p = 0x1.a01a01a01a01ap-13 + r * p;          /* 1/7!  */
    p = 0x1.6c16c16c16c17p-10 + r * p;          /* 1/6!  */
    p = 0x1.1111111111111p-7 + r * p;           /* 1/5!  */
    p = 0x1.5555555555555p-5 + r * p;           /* 1/4!  */
    p = 0x1.5555555555555p-3 + r * p;           /* 1/3!  */
    p = 0.5 + r * p;                            /* 1/2!  */
    p = 1.0 + r * p;
    return 1.0 + r * p;
}

DCM_FN double dcm_exp(double x)
{
    if (x != x) return x;
    if (x > 709.782712893384) return HUGE_VAL;
    if (x < -745.1332191019412) return 0.0;
    const double kd = rint(x * DCM_LOG2E);
    const int k = (int)kd;
    const double r = (x - kd * DCM_LN2_HI) - kd * DCM_LN2_LO;
    const double e = dcm_exp_poly(r);
    if (k > 1023) return (e * dcm_pow2i(k - 1)) * 2.0;
    if (k < -1021) return (e * dcm_pow2i(k + 1000)) * dcm_pow2i(-1000);
    return e * dcm_pow2i(k);
}

/* e^y - 1 for |y| <= 1 (Taylor, degree 20) */
DCM_FN double dcm_expm1_small(double y)
{
    double p = 0x1.e542ba4020225p-62;          /* 1/20! */
    p = 0x1.2f49b46814157p-57 + y * p;          /* 1/19! */
    p = 0x1.6827863b97d97p-53 + y * p;          /* 1/18! */
    p = 0x1.952c77030ad4ap-49 + y * p;          /* 1/17! */
    p = 0x1.ae7f3e733b81fp-45 + y * p;          /* 1/16! */
    p = 0x1.ae7f3e733b81fp-41 + y * p;          /* 1/15! */
    p = 0x1.93974a8c07c9dp-37 + y * p;          /* 1/14! */
    p = 0x1.6124613a86d09p-33 + y * p;          /* 1/13! */
    p = 0x1.1eed8eff8d898p-29 + y * p;          /* 1/12! */
    p = 0x1.ae64567f544e4p-26 + y * p;          /* 1/11! */
    p = 0x1.27e4fb7789f5cp-22 + y * p;          /* 1/10! */
    p = 0x1.71de3a556c734p-19 + y * p;          /* 1/9!  */
    p = 0x1.a01a01a01a01ap-16 + y * p;          /* 1/8!  */
    p = 0x1.a01a01a01a01ap-13 + y * p;          /* 1/7!  */
    p = 0x1.6c16c16c16c17p-10 + y * p;          /* 1/6!  */
    p = 0x1.1111111111111p-7 + y * p;           /* 1/5!  */
    p = 0x1.5555555555555p-5 + y * p;           /* 1/4!  */
    p = 0x1.5555555555555p-3 + y * p;           /* 1/3!  */
    p = 0.5 + y * p;                            /* 1/2!  */
    return y + (y * y) * p;
}

DCM_FN double dcm_tanh(double x)
{
    if (x != x || x == 0.0) return x;
    const double ax = fabs(x);
    double t;
    if (ax > 22.0) {
        t = 1.0;
    } else if (ax >= 0.5) {
        const double e = dcm_exp(2.0 * ax);
        t = 1.0 - 2.0 / (e + 1.0);
    } else {
        const double e = dcm_expm1_small(2.0 * ax);
        t = e / (e + 2.0);
    }
    return copysign(t, x);
}

/* atan(t) - t = -t*Q(t^2) for |t| < 7/16: Q = z/3 - z^2/5 + ... (to t^49) */
DCM_FN double dcm_atan_q(double z)
{
    double q = -0x1.4e5e0a72f0539p-6;           /* -1/49 */
    q = 0x1.5c9882b931057p-6 + z * q;            /* 1/47 */
    q = -0x1.6c16c16c16c17p-6 + z * q;           /* -1/45 */
    q = 0x1.7d05f417d05f4p-6 + z * q;            /* 1/43 */
    q = -0x1.8f9c18f9c18fap-6 + z * q;           /* -1/41 */
    q = 0x1.a41a41a41a41ap-6 + z * q;            /* 1/39 */
    q = -0x1.bacf914c1bad0p-6 + z * q;           /* -1/37 */
    q = 0x1.d41d41d41d41dp-6 + z * q;            /* 1/35 */
    q = -0x1.f07c1f07c1f08p-6 + z * q;           /* -1/33 */
    q = 0x1.0842108421084p-5 + z * q;            /* 1/31 */
    q = -0x1.1a7b9611a7b96p-5 + z * q;           /* -1/29 */
    q = 0x1.2f684bda12f68p-5 + z * q;            /* 1/27 */
    q = -0x1.47ae147ae147bp-5 + z * q;           /* -1/25 */
    q = 0x1.642c8590b2164p-5 + z * q;            /* 1/23 */
    q = -0x1.8618618618618p-5 + z * q;           /* -1/21 */
    q = 0x1.af286bca1af28p-5 + z * q;            /* 1/19 */
    q = -0x1.e1e1e1e1e1e1ep-5 + z * q;           /* -1/17 */
    q = 0x1.1111111111111p-4 + z * q;            /* 1/15 */
    q = -0x1.3b13b13b13b14p-4 + z * q;           /* -1/13 */
    q = 0x1.745d1745d1746p-4 + z * q;            /* 1/11 */
    q = -0x1.c71c71c71c71cp-4 + z * q;           /* -1/9 */
    q = 0x1.2492492492492p-3 + z * q;            /* 1/7 */
    q = -0x1.999999999999ap-3 + z * q;           /* -1/5 */
    q = 0x1.5555555555555p-2 + z * q;            /* 1/3 */
    return z * q;
}

DCM_FN double dcm_atan(double x)
{
    if (x != x || x == 0.0) return x;
    const double ax = fabs(x);
    if (ax >= 0x1p66) return copysign(0x1.921fb54442d18p+0, x);
    double t, hi = 0.0, lo = 0.0;
    int red = 1;
    if (ax < 0.4375) {
        t = ax;
        red = 0;
    } else if (ax < 0.6875) {
        t = (2.0 * ax - 1.0) / (2.0 + ax);
        hi = 0x1.dac670561bb4fp-2;  lo = 0x1.a2b7f222f65e2p-56;   /* atan(1/2) */
    } else if (ax < 1.1875) {
        t = (ax - 1.0) / (ax + 1.0);
        hi = 0x1.921fb54442d18p-1;  lo = 0x1.1a62633145c07p-55;   /* atan(1) */
    } else if (ax < 2.4375) {
        t = (ax - 1.5) / (1.0 + 1.5 * ax);
        hi = 0x1.f730bd281f69bp-1;  lo = 0x1.007887af0cbbdp-56;   /* atan(3/2) */
    } else {
        t = -1.0 / ax;
        hi = 0x1.921fb54442d18p+0;  lo = 0x1.1a62633145c07p-54;   /* pi/2 */
    }
    const double tq = t * dcm_atan_q(t * t);
    const double res = red ? hi - ((tq - lo) - t) : t - tq;
    return copysign(res, x);
}

#define DCM_PIO2_1 0x1.921fb54400000p+0
#define DCM_PIO2_2 0x1.0b4611a600000p-34
#define DCM_PIO2_3 0x1.3198a2e037073p-69
#define DCM_TWO_OVER_PI 0x1.45f306dc9c883p-1

DCM_FN double dcm_tan(double x)
{
    if (x != x || x == 0.0) return x;
    if (x - x != 0.0) return x - x;             /* +-inf -> NaN */
    const double kd = rint(x * DCM_TWO_OVER_PI);
    const double r = ((x - kd * DCM_PIO2_1) - kd * DCM_PIO2_2) - kd * DCM_PIO2_3;
    const double z = r * r;
    /* sin r = r + r z (-1/3! + z (1/5! - ... - z^8/19!)) */
    double sp = -0x1.2f49b46814157p-57;         /* -1/19! */
    sp = 0x1.952c77030ad4ap-49 + z * sp;         /* +1/17! */
    sp = -0x1.ae7f3e733b81fp-41 + z * sp;        /* -1/15! */
    sp = 0x1.6124613a86d09p-33 + z * sp;         /* +1/13! */
    sp = -0x1.ae64567f544e4p-26 + z * sp;        /* -1/11! */
    sp = 0x1.71de3a556c734p-19 + z * sp;         /* +1/9!  */
    sp = -0x1.a01a01a01a01ap-13 + z * sp;        /* -1/7!  */
    sp = 0x1.1111111111111p-7 + z * sp;          /* +1/5!  */
    sp = -0x1.5555555555555p-3 + z * sp;         /* -1/3!  */
    const double sn = r + (r * z) * sp;
    /* cos r = 1 + z (-1/2! + z (1/4! - ... - z^8/18!)) */
    double cp = -0x1.6827863b97d97p-53;         /* -1/18! */
    cp = 0x1.ae7f3e733b81fp-45 + z * cp;         /* +1/16! */
    cp = -0x1.93974a8c07c9dp-37 + z * cp;        /* -1/14! */
    cp = 0x1.1eed8eff8d898p-29 + z * cp;         /* +1/12! */
    cp = -0x1.27e4fb7789f5cp-22 + z * cp;        /* -1/10! */
    cp = 0x1.a01a01a01a01ap-16 + z * cp;         /* +1/8!  */
    cp = -0x1.6c16c16c16c17p-10 + z * cp;        /* -1/6!  */
    cp = 0x1.5555555555555p-5 + z * cp;          /* +1/4!  */
    cp = -0.5 + z * cp;                          /* -1/2!  */
    const double cs = 1.0 + z * cp;
    const int64_t k = (int64_t)kd;
    return (k & 1) ? -cs / sn : sn / cs;
}

/* natural log for x > 0; NaN for x < 0, -inf at 0 */
DCM_FN double dcm_log(double x)
{
    if (x != x) return x;
    if (x < 0.0) return (x - x) / 0.0;
    if (x == 0.0) return -HUGE_VAL;
    if (x - x != 0.0) return x;                 /* +inf */
    uint64_t b = dcm_bits(x);
    int ex = (int)((b >> 52) & 0x7ff);
    int adj = 0;
    if (ex == 0) {                              /* subnormal */
        x = x * 0x1p54;
        b = dcm_bits(x);
        ex = (int)((b >> 52) & 0x7ff);
        adj = -54;
    }
    int e = ex - 1022 + adj;
    double f = dcm_from_bits((b & 0x000fffffffffffffULL) | (1022ULL << 52));   /* [0.5, 1) */
    if (f < 0x1.6a09e667f3bcdp-1) {             /* sqrt(1/2) */
        f = f * 2.0;
        e = e - 1;
    }
    const double s = (f - 1.0) / (f + 1.0);
    const double z = s * s;
    double p = 0x1.642c8590b2164p-5;            /* 1/23 */
    p = 0x1.8618618618618p-5 + z * p;            /* 1/21 */
    p = 0x1.af286bca1af28p-5 + z * p;            /* 1/19 */
    p = 0x1.e1e1e1e1e1e1ep-5 + z * p;            /* 1/17 */
    p = 0x1.1111111111111p-4 + z * p;            /* 1/15 */
    p = 0x1.3b13b13b13b14p-4 + z * p;            /* 1/13 */
    p = 0x1.745d1745d1746p-4 + z * p;            /* 1/11 */
    p = 0x1.c71c71c71c71cp-4 + z * p;            /* 1/9 */
    p = 0x1.2492492492492p-3 + z * p;            /* 1/7 */
    p = 0x1.999999999999ap-3 + z * p;            /* 1/5 */
    p = 0x1.5555555555555p-2 + z * p;            /* 1/3 */
    const double s2 = 2.0 * s;
    const double ed = (double)e;
    return ed * DCM_LN2_HI + ((ed * DCM_LN2_LO + s2 * (z * p)) + s2);
}

/* x^y for x >= 0 (the optional segregation and Manning terms only) */
DCM_FN double dcm_pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (x == 1.0) return 1.0;
    if (x == 0.0) return y > 0.0 ? 0.0 : HUGE_VAL;
    return dcm_exp(y * dcm_log(x));
}

#endif
