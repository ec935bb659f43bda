/*
 * plom_psr.h -- the discrete samplers of the Poisson-system-with-stochastic-rates step
 * (implementation "psr", step_psr, C/templates/core_template.c:104-209), shared by the GPU
 * kernels and the CPU oracle (compiled with -ffp-contract=off on both sides, so a particle
 * draws the same numbers on both).
 *
 * Every draw consumes 32-bit words of the particle's xoshiro128** stream (plom_rng.h) in
 * a fixed order; the arithmetic uses this library's exp and sqrt (plom_math.h) and the
 * table logarithm plom_log_tab (plom_rng.h: no division, no branch) on the stream's log
 * table ns->ltab.  Each rejection test takes one logarithm of a product instead of a sum
 * of logarithms.
 *
 *   plom_unif53        (a >> 6) 2^27 + (b >> 5) + 1/2, times 2^-53: a uniform in (0, 1)
 *                      from two words (the inversion and multiplication methods, whose
 *                      tails need the resolution).
 *   plom_unif32        (w + 1/2) 2^-32 from one word: the uniforms of the rejection
 *                      attempts (transformed rejection, Marsaglia-Tsang acceptance).
 *   plom_lfact         log k!: exact table below 30, Stirling's series above (error
 *                      below 1e-14); the stream carries both tables (plom_psr_tables).
 *                      plom_lfact2 sums two of them with one division for both.
 * The transformed-rejection attempts take a2/us and us^2/(a + b us^2) from one reciprocal,
 * and the setups pair their divisions the same way: fewer FP64 division sequences, each
 * value within a few ulps of the separate quotient.
 *   plom_gamma_draw    Gamma(a, 1): Marsaglia and Tsang (2000, ACM TOMS 26:363-372) for
 *                      a >= 1; Gamma(a + 1) U^(1/a) below.
 *   plom_poisson_draw  Poisson(mu): the multiplication method (Knuth, TAOCP vol. 2,
 *                      3.4.1) for mu < 10; Hoermann's transformed rejection with squeeze
 *                      PTRS (1993, Insurance: Math. Econ. 12:39-45) above.
 *   plom_binom_draw    Binomial(n, p), p <= 1/2 by symmetry: inversion by sequential
 *                      search (Kachitvichyanukul and Schmeiser 1988, BINV) for n p < 10;
 *                      Hoermann's BTRS (1993, J. Stat. Comput. Simul. 46:101-110) above.
 *   plom_multinomial   the sequential binomials of plom_ran_multinomial
 *                      (C/core/prediction_util.c:323-343).
 *
 * The rejection loops are bounded (PLOM_PSR_MAX_TRY attempts; the probability of reaching
 * the bound is below 1e-60 for every parameter) and then return the distribution's mode,
 * so that every GPU lane leaves the loop.
 */
#ifndef PLOM_PSR_H
#define PLOM_PSR_H

#include "plom_math.h"
#include "plom_rng.h"

#define PLOM_PSR_MAX_TRY 64

/* the samplers are real calls on the device: a step holds dozens of draws, and inlining
 * the rejection loops at every one of them multiplies the code size */
#if defined(__HIPCC__)
#define PLOM_HD_CALL static __host__ __device__ __attribute__((noinline))
#else
#define PLOM_HD_CALL static
#endif

PLOM_HD double plom_floor(double x) { return __builtin_floor(x); }
#define PLOM_LN_SQRT_2PI 0.91893853320467274178

PLOM_HD double plom_unif53(plom_nstream64 *ns)
{
    const uint32_t a = plom_xoshiro_next(&ns->r) >> 6;
    const uint32_t b = plom_xoshiro_next(&ns->r) >> 5;
    return ((double)a * 134217728.0 + (double)b + 0.5) * 0x1p-53;
}

PLOM_HD double plom_unif32(plom_nstream64 *ns)
{
    return plom_fma((double)plom_xoshiro_next(&ns->r), 0x1p-32, 0x1p-33); /* exact */
}

/* log k! for k = 0..29 (50-digit values rounded to nearest) */
#define PLOM_LFACT_TAB { \
    0.0, 0.0, 0.6931471805599453, 1.791759469228055, 3.1780538303479458, \
    4.787491742782046, 6.579251212010101, 8.525161361065415, 10.60460290274525, 12.801827480081469, \
    15.104412573075516, 17.502307845873887, 19.987214495661885, 22.552163853123425, 25.19122118273868, \
    27.89927138384089, 30.671860106080672, 33.50507345013689, 36.39544520803305, 39.339884187199495, \
    42.335616460753485, 45.38013889847691, 48.47118135183523, 51.60667556776438, 54.78472939811232, \
    58.00360522298052, 61.261701761002, 64.55753862700634, 67.88974313718154, 71.25703896716801 \
}

#if defined(__HIPCC__)
static __constant__ const double plom_lfact_tab[30] = PLOM_LFACT_TAB;
#else
static const double plom_lfact_tab[30] = PLOM_LFACT_TAB;
#endif

/* log k! for an integer k >= 0: the table (lftab: plom_lfact_tab or a copy) below 30,
 * Stirling's series above; both are evaluated, without a branch */
/* log k! with x = k + 1 and r = 1/x given */
PLOM_HD double plom_lfact_r(double k, double x, double r, const double *ltab, const double *lftab)
{
    const double r2 = r * r;
    /* ln Gamma(x) = (x - 1/2) ln x - x + ln sqrt(2 pi) + 1/(12x) - 1/(360x^3) + 1/(1260x^5) */
    const double st = (x - 0.5) * plom_log_tab(x, ltab) - x + PLOM_LN_SQRT_2PI
                      + r * (1.0 / 12.0 - r2 * (1.0 / 360.0 - r2 * (1.0 / 1260.0)));
    const int small = k < 30.0;
    const double tv = lftab[small ? (int)k : 0];
    return small ? tv : st;
}

PLOM_HD double plom_lfact(double k, const double *ltab, const double *lftab)
{
    const double x = k + 1.0;
    return plom_lfact_r(k, x, 1.0 / x, ltab, lftab);
}

/* log k1! + log k2! with one division for both series: 1/x1 = x2 R and 1/x2 = x1 R from
 * R = 1/(x1 x2) (x1 x2 exact below 2^53; a few ulps either way, far below what the
 * 1/(12x) term needs) */
PLOM_HD double plom_lfact2(double k1, double k2, const double *ltab, const double *lftab)
{
    const double x1 = k1 + 1.0, x2 = k2 + 1.0;
    const double R = 1.0 / (x1 * x2);
    return plom_lfact_r(k1, x1, x2 * R, ltab, lftab) + plom_lfact_r(k2, x2, x1 * R, ltab, lftab);
}

/* the stream's psr tables: the log table it holds and the log k! table (the host's) */
PLOM_HD void plom_psr_tables(plom_nstream64 *ns, const double *ltab, const double *lftab)
{
    ns->ltab = ltab;
    ns->lftab = lftab;
}

/* Gamma(a, 1), a > 0 */
PLOM_HD_CALL double plom_gamma_draw(plom_nstream64 *ns, double a)
{
    double boost = 1.0;
    if (a < 1.0) {
        /* Gamma(a) = Gamma(a + 1) U^(1/a) */
        boost = plom_exp(plom_log_tab(plom_unif53(ns), ns->ltab) / a);
        a = a + 1.0;
    }
    const double d = a - 1.0 / 3.0;
    const double c = 1.0 / plom_sqrt(9.0 * d);
    for (int t = 0; t < PLOM_PSR_MAX_TRY; t++) {
        double z, z1;
        plom_normal_pair64(ns, &z, &z1);
        const double v1 = 1.0 + c * z;
        if (v1 <= 0.0) continue;
        const double v = v1 * v1 * v1;
        const double u = plom_unif32(ns);
        const double z2 = z * z;
        if (u < 1.0 - 0.0331 * (z2 * z2)) return (d * v) * boost;
        /* v underflows to 0 only for v1 < 1e-102, where the test fails either way */
        if (plom_log_tab(u, ns->ltab) < 0.5 * z2 + d * (1.0 - v + plom_log_tab(v, ns->ltab))) return (d * v) * boost;
    }
    return d * boost;
}

/* Poisson(mu), mu >= 0, as a double holding an integer */
PLOM_HD_CALL double plom_poisson_draw(plom_nstream64 *ns, double mu)
{
    if (!(mu > 0.0)) return 0.0;
    if (mu < 10.0) {
        const double L = plom_exp(-mu);
        double p = plom_unif53(ns);
        double k = 0.0;
        /* P(k > 200) < 1e-150 for mu < 10 */
        while (p > L && k < 200.0) {
            k = k + 1.0;
            p = p * plom_unif53(ns);
        }
        return k;
    }
    const double slam = plom_sqrt(mu);
    const double loglam = plom_log_tab(mu, ns->ltab);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    /* 1/(b - 3.4) and 1/(b - 2) from one division */
    const double bm34 = b - 3.4, bm2 = b - 2.0;
    const double Rb = 1.0 / (bm34 * bm2);
    const double invalpha = 1.1239 + 1.1328 * (bm2 * Rb);
    const double vr = 0.9277 - 3.6224 * (bm34 * Rb);
    const double a2 = 2.0 * a;
    for (int t = 0; t < PLOM_PSR_MAX_TRY; t++) {
        const double U = plom_unif32(ns) - 0.5;
        const double V = plom_unif32(ns);
        const double us = 0.5 - plom_fabs(U);
        /* a2 / us and us^2 / (a + b us^2) from one division: R = 1 / (us (a + b us^2)) */
        const double us2 = us * us;
        const double den = plom_fma(b, us2, a);
        const double R = 1.0 / (us * den);
        const double k = plom_floor((a2 * (den * R) + b) * U + mu + 0.43);
        /* the quick acceptance, the early rejection and the full test, evaluated together
         * (no branch: a wave takes the full test whenever one of its lanes needs it), the
         * full test at k clamped to >= 0 where it is not used */
        const double kk = k < 0.0 ? 0.0 : k;
        /* log(V invalpha / (a / us^2 + b)) */
        const double lhs = plom_log_tab(V * invalpha * (us2 * us * R), ns->ltab);
        const double rhs = -mu + kk * loglam - plom_lfact(kk, ns->ltab, ns->lftab);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (lhs <= rhs) return k;
    }
    return plom_floor(mu);
}

/* Binomial(n, p) for p <= 1/2 */
PLOM_HD double plom_binom_half(plom_nstream64 *ns, double n, double p)
{
    const double q = 1.0 - p;
    const double np = n * p;
    if (np < 10.0) {
        /* inversion: r = P(X = x) by the recursion r(x+1) = r(x) (a/(x+1) - s) */
        const double s = p / q;
        const double a = (n + 1.0) * s;
        /* q^n = exp(n log(1 - p)); log(1 - p) by its series when p is small */
        double l1p;
        if (p < 1e-4) {
            const double p2 = p * p;
            l1p = -(p + p2 * (0.5 + p * (1.0 / 3.0 + p * 0.25)));
        } else {
            l1p = plom_log_tab(q, ns->ltab);
        }
        double r = plom_exp(n * l1p);
        double u = plom_unif53(ns);
        double x = 0.0;
        /* P(X > 110) < 1e-60 for n p < 10 */
        while (u > r && x < n && x < 110.0) {
            u = u - r;
            x = x + 1.0;
            r = r * (a / x - s);
        }
        return x;
    }
    const double spq = plom_sqrt(np * q);
    const double b = 1.15 + 2.53 * spq;
    const double a = -0.0873 + 0.0248 * b + 0.01 * p;
    const double c = np + 0.5;
    /* 1/b and p/q from one division: Rq = 1 / (b q) */
    const double Rq = 1.0 / (b * q);
    const double rb = q * Rq;
    const double vr = 0.92 - 4.2 * rb;
    const double alpha = (2.83 + 5.1 * rb) * spq;
    const double a2 = 2.0 * a;
    const double lpq = plom_log_tab(p * (b * Rq), ns->ltab);
    const double m = plom_floor((n + 1.0) * p);
    const double h = plom_lfact2(m, n - m, ns->ltab, ns->lftab);
    for (int t = 0; t < PLOM_PSR_MAX_TRY; t++) {
        const double u = plom_unif32(ns) - 0.5;
        const double v = plom_unif32(ns);
        const double us = 0.5 - plom_fabs(u);
        /* a2 / us and us^2 / (a + b us^2) from one division: R = 1 / (us (a + b us^2)) */
        const double us2 = us * us;
        const double den = plom_fma(b, us2, a);
        const double R = 1.0 / (us * den);
        const double k = plom_floor((a2 * (den * R) + b) * u + c);
        /* the range check, the quick acceptance and the full test, evaluated together (no
         * branch: a wave takes the full test whenever one of its lanes needs it), the full
         * test at k clamped into [0, n] where it is not used */
        const double kk = k < 0.0 ? 0.0 : (k > n ? n : k);
        /* log(v alpha / (a / us^2 + b)) */
        const double lhs = plom_log_tab(v * alpha * (us2 * us * R), ns->ltab);
        const double rhs = h - plom_lfact2(kk, n - kk, ns->ltab, ns->lftab) + (kk - m) * lpq;
        if (k < 0.0 || k > n) continue;
        if (us >= 0.07 && v <= vr) return k;
        if (lhs <= rhs) return k;
    }
    return m;
}

/* Binomial(n, p), n a non-negative integer held in a double */
PLOM_HD_CALL double plom_binom_draw(plom_nstream64 *ns, double n, double p)
{
    if (!(n > 0.0) || !(p > 0.0)) return 0.0;
    if (p >= 1.0) return n;
    /* one inlined copy of the sampler for both halves */
    const int hi = p > 0.5;
    const double x = plom_binom_half(ns, n, hi ? 1.0 - p : p);
    return hi ? n - x : x;
}

/* n[0..K) ~ Multinomial(N, p[0..K)), p summing to 1: the sequential binomials of
 * plom_ran_multinomial (prediction_util.c:323-343) */
PLOM_HD_CALL void plom_multinomial(plom_nstream64 *ns, int K, double N, const double *p, double *n)
{
    double sum_p = 0.0, sum_n = 0.0;
    for (int k = 0; k < K; k++) {
        if (p[k] > 0.0) n[k] = plom_binom_draw(ns, N - sum_n, p[k] / (1.0 - sum_p));
        else n[k] = 0.0;
        sum_p += p[k];
        sum_n += n[k];
    }
}

#endif
