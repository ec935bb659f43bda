/*
 * plom_math.h -- scalar arithmetic shared, text for text, by the HIP kernels (gfx950)
 * and the C oracle.
 *
 * Every routine here uses only IEEE-754 +, -, *, /, sqrt, fma and integer/bit
 * operations, so a host build (gcc, -ffp-contract=off) and a device build (hipcc,
 * -ffp-contract=off) produce bit-identical doubles.  That is what makes per-particle
 * log-weights, ancestors and the log-likelihood reproducible between the GPU filter and
 * its CPU oracle.  libm and the device math library are deliberately not used: their
 * last-bit behaviour differs between the two sides.
 *
 * Methods (all from the open literature, coefficients computed for this file with
 * mpmath; see gen_phi_table.py for the normal CDF tables):
 *   plom_exp   Cody & Waite (1980) reduction x = k ln2 + r, |r| <= ln2/2, with a two-part
 *              ln2, then the degree-13 Taylor polynomial of e^r (truncation < 5e-18).
 *   plom_log   x = 2^e m, m in [sqrt(2)/2, sqrt(2)), s = f/(2+f), f = m-1;
 *              log(1+f) = f - (f^2/2 - s (f^2/2 + R(s^2))), R the atanh series
 *              sum_{k>=1} 2 s^{2k}/(2k+1) (11 terms, truncation < 1e-18).
 *   plom_sin/cos  three-part Cody-Waite reduction by pi/2, Taylor kernels on |r| <= pi/4.
 *   plom_Phi   standard normal CDF: Chebyshev interpolants of (Phi(x)-1/2)/x in x^2 on
 *              |x|<=1 and of the Mills ratio t*Phi(-t)/phi(t) in 1/t on four bands,
 *              evaluated by the Clenshaw recurrence, with an exactly split t^2 inside
 *              exp(-t^2/2).
 * Accuracy (checked against libm/scipy/mpmath in tests/test_math_rng.py): exp, log,
 * sin, cos within 2 ulp; Phi within 1e-15 relative for Phi >= 1e-300.
 */
#ifndef PLOM_MATH_H
#define PLOM_MATH_H

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define PLOM_HD static __host__ __device__ inline __attribute__((always_inline))
#else
#define PLOM_HD static inline
#endif

#include "plom_phi_coeffs.h"

#define PLOM_INF (__builtin_inf())
#define PLOM_NAN (__builtin_nan(""))

/* ---------------------------------------------------------------- bit helpers */
PLOM_HD uint64_t plom_d2u(double x) { union { double d; uint64_t u; } v; v.d = x; return v.u; }
PLOM_HD double plom_u2d(uint64_t u) { union { double d; uint64_t u; } v; v.u = u; return v.d; }
PLOM_HD double plom_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
PLOM_HD double plom_sqrt(double x) { return __builtin_sqrt(x); }

/* IEEE sqrt of the Box-Muller radius at a lower cost on the device.  gfx950 has no
 * correctly rounded f64 sqrt instruction; the compiler refines v_rsq_f64 by
 * Newton-Raphson steps (the core below) and wraps that in a range scaling (inputs below
 * 2^-767) and a pass-through of 0 and +inf.  For x in [2^-767, 2^1023] the scaling is by
 * 2^0, so the core alone is the same instruction sequence with the same correctly rounded
 * result.  plom_sqrt_mid takes only such x: -2 log u lies in [2.3e-10, 46].  The host
 * computes IEEE sqrt. */
#if defined(__HIP_DEVICE_COMPILE__)
PLOM_HD double plom_sqrt_mid(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = 0.5 * y;
    const double r = plom_fma(-h, g, 0.5);
    g = plom_fma(g, r, g);
    h = plom_fma(h, r, h);
    double d = plom_fma(-g, g, x);
    g = plom_fma(d, h, g);
    d = plom_fma(-g, g, x);
    return plom_fma(d, h, g);
}
#else
PLOM_HD double plom_sqrt_mid(double x) { return __builtin_sqrt(x); }
#endif

/* The square root of a demographic noise term, sqrt(rate x state) of the Euler-Maruyama
 * step (core_template.c:292-293), as this build defines it on both sides:
 *   0          for +0 <= x < 2^-767 (a bit-pattern compare: the +0 and tiny range),
 *   sqrt(x)    (IEEE) for 2^-767 <= x < +inf,
 *   NaN        otherwise (x < 0, -0, +inf, NaN).
 * It differs from IEEE sqrt only below 2^-767 (where sqrt is < 2^-383, far below an ulp of
 * any state it is added to) and at -0 and +inf.  The device computes the Newton core of
 * plom_sqrt_mid, exact on [2^-767, +inf) and NaN on the last class, and selects 0 for the
 * first, so the range scaling of the IEEE expansion is not needed. */
#define PLOM_SQRT_RATE_LO 0x1000000000000000ULL /* the bits of 2^-767 */
#if defined(__HIP_DEVICE_COMPILE__)
PLOM_HD double plom_sqrt_rate(double x)
{
    const double r = plom_sqrt_mid(x);
    return plom_d2u(x) < PLOM_SQRT_RATE_LO ? 0.0 : r;
}
#else
PLOM_HD double plom_sqrt_rate(double x)
{
    if (plom_d2u(x) < PLOM_SQRT_RATE_LO) return 0.0;
    if (x > 0.0 && x < PLOM_INF) return __builtin_sqrt(x);
    return PLOM_NAN;
}
#endif
/* 1 / c (IEEE), for a divisor shared by many quotients */
PLOM_HD double plom_recip_uniform(double c) { return 1.0 / c; }

/* x / c from rc = RN(1/c), in three operations instead of a division: q = RN(x rc), the
 * residual e = RN(x - q c) (one fma, exact when q is within an ulp of x / c), and the
 * corrected quotient RN(q + e rc) (Markstein's correction step), with the sign of q, which
 * is the sign of x / c also when the quotient is zero.  For finite x and c with no
 * underflow the result is the IEEE quotient x / c except in rare cases close to a rounding
 * midpoint, where it can be one ulp away (the correction step is exact only for a faithful
 * q; see Brisebarre, Muller and Raina, "Accelerating correctly rounded floating-point
 * division when the divisor is known in advance", IEEE Trans. Computers 53(8), 2004).
 * x = +-inf gives NaN.  The same operations on the device and the host. */
PLOM_HD double plom_div_rc(double x, double c, double rc)
{
    const double q = x * rc;
    const double e = plom_fma(-q, c, x);
    const double r = plom_fma(e, rc, q);
    return plom_u2d((plom_d2u(r) & 0x7fffffffffffffffULL) | (plom_d2u(q) & 0x8000000000000000ULL));
}

PLOM_HD double plom_fabs(double x) { return plom_u2d(plom_d2u(x) & 0x7fffffffffffffffULL); }
PLOM_HD int plom_isnan(double x) { return x != x; }
/* 2^k for k in [-1022, 1023] */
PLOM_HD double plom_pow2i(int k) { return plom_u2d((uint64_t)(k + 1023) << 52); }
/* round-to-nearest-even integer value of x, |x| < 2^51 */
PLOM_HD double plom_rint(double x)
{
    const double magic = 0x1.8p52;
    double y = x + magic;
    return y - magic;
}

/* ---------------------------------------------------------------- exp */
#define PLOM_LN2_HI 0x1.62e42fefa2000p-1
#define PLOM_LN2_LO 0x1.9ef35793c7673p-41
#define PLOM_INV_LN2 0x1.71547652b82fep+0

PLOM_HD double plom_exp(double x)
{
    if (plom_isnan(x)) return x;
    if (x > 709.782712893384) return PLOM_INF;
    if (x < -745.1332191019412) return 0.0;
    double kd = plom_rint(x * PLOM_INV_LN2);
    int k = (int)kd;
    double r = x - kd * PLOM_LN2_HI;
    r = r - kd * PLOM_LN2_LO;
    double p = 0x1.6124613a86d09p-33;
    p = plom_fma(p, r, 0x1.1eed8eff8d898p-29);
    p = plom_fma(p, r, 0x1.ae64567f544e4p-26);
    p = plom_fma(p, r, 0x1.27e4fb7789f5cp-22);
    p = plom_fma(p, r, 0x1.71de3a556c734p-19);
    p = plom_fma(p, r, 0x1.a01a01a01a01ap-16);
    p = plom_fma(p, r, 0x1.a01a01a01a01ap-13);
    p = plom_fma(p, r, 0x1.6c16c16c16c17p-10);
    p = plom_fma(p, r, 0x1.1111111111111p-7);
    p = plom_fma(p, r, 0x1.5555555555555p-5);
    p = plom_fma(p, r, 0x1.5555555555555p-3);
    p = plom_fma(p, r, 0x1.0000000000000p-1);
    p = plom_fma(p, r, 1.0);
    p = plom_fma(p, r, 1.0);
    /* scale by 2^k in two exact steps (the second may round into the subnormals) */
    int k1 = k / 2;
    int k2 = k - k1;
    return (p * plom_pow2i(k1)) * plom_pow2i(k2);
}

/* ---------------------------------------------------------------- log */
/* log of a positive normal finite x (the core of plom_log, also used directly by the
 * Box-Muller transform whose argument is always in [2^-33, 1)) */
PLOM_HD double plom_log_core(double x, int e)
{
    uint64_t u = plom_d2u(x);
    e += (int)(u >> 52) - 1023;
    double m = plom_u2d((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    const int big = m > 0x1.6a09e667f3bcdp+0;
    m = big ? m * 0.5 : m;
    e += big;
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double R = 0x1.642c8590b2164p-4;
    R = plom_fma(R, z, 0x1.8618618618618p-4);
    R = plom_fma(R, z, 0x1.af286bca1af28p-4);
    R = plom_fma(R, z, 0x1.e1e1e1e1e1e1ep-4);
    R = plom_fma(R, z, 0x1.1111111111111p-3);
    R = plom_fma(R, z, 0x1.3b13b13b13b14p-3);
    R = plom_fma(R, z, 0x1.745d1745d1746p-3);
    R = plom_fma(R, z, 0x1.c71c71c71c71cp-3);
    R = plom_fma(R, z, 0x1.2492492492492p-2);
    R = plom_fma(R, z, 0x1.999999999999ap-2);
    R = plom_fma(R, z, 0x1.5555555555555p-1);
    R = R * z;
    double hfsq = 0.5 * f * f;
    double lm = f - (hfsq - s * (hfsq + R));
    double de = (double)e;
    return de * PLOM_LN2_HI + (de * PLOM_LN2_LO + lm);
}

PLOM_HD double plom_log(double x)
{
    if (plom_isnan(x)) return x;
    if (x < 0.0) return PLOM_NAN;
    if (x == 0.0) return -PLOM_INF;
    if (x == PLOM_INF) return x;
    if (x < 0x1p-1022) return plom_log_core(x * 0x1p54, -54);
    return plom_log_core(x, 0);
}

/* ---------------------------------------------------------------- sin / cos */
/* Taylor kernels on |r| <= pi/4 */
PLOM_HD double plom_sin_kernel(double r)
{
    double z = r * r;
    double p = 0x1.952c77030ad4ap-49;
    p = plom_fma(p, z, -0x1.ae7f3e733b81fp-41);
    p = plom_fma(p, z, 0x1.6124613a86d09p-33);
    p = plom_fma(p, z, -0x1.ae64567f544e4p-26);
    p = plom_fma(p, z, 0x1.71de3a556c734p-19);
    p = plom_fma(p, z, -0x1.a01a01a01a01ap-13);
    p = plom_fma(p, z, 0x1.1111111111111p-7);
    p = plom_fma(p, z, -0x1.5555555555555p-3);
    return plom_fma(r * z, p, r);
}

PLOM_HD double plom_cos_kernel(double r)
{
    double z = r * r;
    double p = -0x1.6827863b97d97p-53;
    p = plom_fma(p, z, 0x1.ae7f3e733b81fp-45);
    p = plom_fma(p, z, -0x1.93974a8c07c9dp-37);
    p = plom_fma(p, z, 0x1.1eed8eff8d898p-29);
    p = plom_fma(p, z, -0x1.27e4fb7789f5cp-22);
    p = plom_fma(p, z, 0x1.a01a01a01a01ap-16);
    p = plom_fma(p, z, -0x1.6c16c16c16c17p-10);
    p = plom_fma(p, z, 0x1.5555555555555p-5);
    double hz = 0.5 * z;
    double w = 1.0 - hz;
    /* 1 - z/2 + z^2 p, with the rounding of 1 - z/2 carried */
    return w + (((1.0 - w) - hz) + z * (z * p));
}

/* sin/cos of r + q pi/2 from sin r, cos r without branches: odd q swaps the pair,
 * q in {2,3} negates sin, q in {1,2} negates cos (sign-bit flips are exact) */
PLOM_HD void plom_quadrant(int q, double sr, double cr, double *s, double *c)
{
    const int odd = q & 1;
    double a = odd ? cr : sr;
    double b = odd ? sr : cr;
    *s = plom_u2d(plom_d2u(a) ^ ((uint64_t)(q & 2) << 62));
    *c = plom_u2d(plom_d2u(b) ^ ((uint64_t)((q + 1) & 2) << 62));
}

#define PLOM_PIO2_1 0x1.921fb54400000p+0
#define PLOM_PIO2_2 0x1.0b4611a600000p-34
#define PLOM_PIO2_3 0x1.3198a2e037073p-69
#define PLOM_2_OVER_PI 0x1.45f306dc9c883p-1

/* sin and cos of x for |x| < 2^20 (ample for forcing terms in days) */
PLOM_HD void plom_sincos(double x, double *s, double *c)
{
    /* outside |x| <= 2^20 the quadrant count would overflow an int (undefined and
     * different between x86 and gfx950): NaN on both sides instead, without a branch */
    const int ok = plom_fabs(x) <= 0x1p20;
    const double xs = ok ? x : 0.0;
    double kd = plom_rint(xs * PLOM_2_OVER_PI);
    double r = xs - kd * PLOM_PIO2_1;
    r = r - kd * PLOM_PIO2_2;
    r = r - kd * PLOM_PIO2_3;
    plom_quadrant(((int)kd) & 3, plom_sin_kernel(r), plom_cos_kernel(r), s, c);
    *s = ok ? *s : PLOM_NAN;
    *c = ok ? *c : PLOM_NAN;
}

PLOM_HD double plom_sin(double x) { double s, c; plom_sincos(x, &s, &c); return s; }
PLOM_HD double plom_cos(double x) { double s, c; plom_sincos(x, &s, &c); return c; }

/* ---------------------------------------------------------------- powers */
/* x^n for a small integer exponent, by repeated products in a fixed order */
PLOM_HD double plom_powi(double x, int n)
{
    int neg = n < 0;
    unsigned int m = neg ? (unsigned int)(-n) : (unsigned int)n;
    double r = 1.0, b = x;
    while (m) {
        if (m & 1u) r = r * b;
        m >>= 1;
        if (m) b = b * b;
    }
    return neg ? 1.0 / r : r;
}

PLOM_HD double plom_pow(double x, double y)
{
    if (y == plom_rint(y) && plom_fabs(y) <= 64.0) return plom_powi(x, (int)y);
    if (x == 0.0) return y > 0.0 ? 0.0 : PLOM_INF;
    return plom_exp(y * plom_log(x));
}

/* ---------------------------------------------------------------- normal CDF */
PLOM_HD double plom_clenshaw(const double *c, int deg, double z)
{
    double z2 = 2.0 * z;
    double b1 = 0.0, b2 = 0.0;
    for (int k = deg; k >= 1; k--) {
        double b0 = plom_fma(z2, b1, c[k] - b2);
        b2 = b1;
        b1 = b0;
    }
    return plom_fma(z, b1, c[0] - b2);
}

#define PLOM_INV_SQRT_2PI 0x1.9884533d43651p-2

/* Phi(-t) for t >= 1 */
PLOM_HD double plom_Phi_lower_tail(double t)
{
    if (t > 40.0) return 0.0;
    double s = 1.0 / t;
    double h;
    if (t <= 2.0) {
        const double c[] = PLOM_PHI_MILLS_1_2_COEFFS;
        h = plom_clenshaw(c, PLOM_PHI_MILLS_1_2_DEG, (2.0 * s - 1.5) * 2.0);
    } else if (t <= 4.0) {
        const double c[] = PLOM_PHI_MILLS_2_4_COEFFS;
        h = plom_clenshaw(c, PLOM_PHI_MILLS_2_4_DEG, (2.0 * s - 0.75) * 4.0);
    } else if (t <= 8.0) {
        const double c[] = PLOM_PHI_MILLS_4_8_COEFFS;
        h = plom_clenshaw(c, PLOM_PHI_MILLS_4_8_DEG, (2.0 * s - 0.375) * 8.0);
    } else {
        const double c[] = PLOM_PHI_MILLS_8_40_COEFFS;
        h = plom_clenshaw(c, PLOM_PHI_MILLS_8_40_DEG, (2.0 * s - 0.15) * 10.0);
    }
    /* t^2 = p + e exactly; exp(-t^2/2) = exp(-p/2) (1 - e/2) to O(e^2) */
    double p = t * t;
    double e = plom_fma(t, t, -p);
    double g = plom_exp(-0.5 * p);
    g = g - g * (0.5 * e);
    return g * (h * s) * PLOM_INV_SQRT_2PI;
}

PLOM_HD double plom_Phi(double x)
{
    if (plom_isnan(x)) return x;
    double ax = plom_fabs(x);
    if (ax <= 1.0) {
        const double c[] = PLOM_PHI_CENTRAL_COEFFS;
        double g = plom_clenshaw(c, PLOM_PHI_CENTRAL_DEG, 2.0 * (x * x) - 1.0);
        return 0.5 + x * g;
    }
    double q = plom_Phi_lower_tail(ax);
    return x < 0.0 ? q : 1.0 - q;
}

/* Gaussian CDF with standard deviation sd, the role of gsl_cdf_gaussian_P(x, sd):
 * sd == 0 gives the step function (0, or 1 for x >= 0). */
PLOM_HD double plom_cdf_gaussian_P(double x, double sd)
{
    return plom_Phi(x / sd);
}

#endif /* PLOM_MATH_H */
