/*
 * plom_rng.h -- counter-keyed random streams shared by the HIP kernels and the C oracle.
 *
 * Replaces the reference's one sequential GSL stream per worker thread
 * (model_builder/C/core/build.c:965-971 seeds 2+id(+thread), :1019-1030 picks mt19937
 * or ranlxs0), whose draws depend on the thread count and on which thread owns which
 * particle.  Here every draw is a pure function of
 *     (seed, stream id, observation interval n, particle j, purpose)
 * so a filter gives the same result whatever the tiling, the batch it runs in, or the
 * other requests beside it, and a lineage can be replayed particle by particle.
 *
 *   Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random numbers: as easy
 *   as 1, 2, 3"), key = (seed, stream id), counter = (j, n, purpose, 0), gives 128
 *   bits per (n, j, purpose).
 *   PURPOSE_PROPAGATE seeds a per-(n, j) xoshiro128** generator (Blackman & Vigna 2018)
 *   that feeds, in order, every Gaussian draw of that particle over interval n: for each
 *   Euler substep, for each population, the demographic then environmental noises, the
 *   order of core_template.c:258-284, then the drift noises (drift.c:39).
 *   PURPOSE_RESAMPLE gives the one systematic-resampling uniform of interval n
 *   (the role of gsl_ran_flat(calc[0]) in smc.c:100).
 *   PURPOSE_FULL gives the uniform that picks the final particle of a --full lineage.
 *   PURPOSE_OBS gives the Gaussians of the observations drawn along a lineage.
 *
 * Uniforms are (x + 1/2) 2^-32 for a 32-bit x: never 0 or 1, exact in double.
 * Gaussians of the stream (the Euler noise): one per 32-bit word, the inverse normal CDF
 * of the word's uniform, evaluated on piecewise polynomials (plom_icdf_normal, table
 * plom_icdf_table.h: no division, no branch, no transcendental); the generated step code
 * draws them two at a time in draw order and drops the second value of a pair left open at
 * the end of an Euler substep, so the pairing is fixed at code-generation time.  The
 * Gaussians drawn from Philox words outside the stream (the MIF random walk, the drawn
 * observations of a lineage) are Box-Muller pairs on two words with small lookup tables
 * (plom_bm_m2log, plom_bm_sincos) and IEEE sqrt.  The FP32 stream (device only) uses the
 * gfx950 transcendental instructions.
 */
#ifndef PLOM_RNG_H
#define PLOM_RNG_H

#include "plom_math.h"
#include "plom_bm_tables.h"
#include "plom_icdf_table.h"

#define PLOM_PURPOSE_PROPAGATE 0u
#define PLOM_PURPOSE_RESAMPLE 1u
#define PLOM_PURPOSE_FULL 2u
#define PLOM_PURPOSE_OBS 3u

PLOM_HD uint32_t plom_mulhi32(uint32_t a, uint32_t b)
{
    return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

PLOM_HD void plom_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t hi0 = plom_mulhi32(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        uint32_t hi1 = plom_mulhi32(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 128 bits for (seed, stream, n, j, purpose) */
PLOM_HD void plom_draw128(uint32_t seed, uint32_t stream, uint32_t n, uint32_t j,
                          uint32_t purpose, uint32_t out[4])
{
    uint32_t ctr[4] = { j, n, purpose, 0u };
    uint32_t key[2] = { seed, stream };
    plom_philox4x32_10(ctr, key, out);
}

typedef struct { uint32_t s0, s1, s2, s3; } plom_xoshiro128;

PLOM_HD uint32_t plom_rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

/* xoshiro128** (Blackman and Vigna): the output rotl(5 s1, 7) 9, then the state update */
PLOM_HD uint32_t plom_xoshiro_next(plom_xoshiro128 *r)
{
    uint32_t result = plom_rotl32(r->s1 * 5u, 7) * 9u;
    uint32_t t = r->s1 << 9;
    r->s2 ^= r->s0;
    r->s3 ^= r->s1;
    r->s1 ^= r->s2;
    r->s0 ^= r->s3;
    r->s2 ^= t;
    r->s3 = plom_rotl32(r->s3, 11);
    return result;
}

PLOM_HD void plom_xoshiro_seed(plom_xoshiro128 *r, const uint32_t w[4])
{
    r->s0 = w[0]; r->s1 = w[1]; r->s2 = w[2]; r->s3 = w[3];
    if ((w[0] | w[1] | w[2] | w[3]) == 0u) r->s0 = 1u; /* the all-zero state is a fixed point */
}

/* ---------------------------------------------------------------- FP64 Gaussians */
/* The lookup tables of the transform (see gen_bm_tables.py).  Kernels copy them to LDS
 * and hand the stream a pointer to that copy; the host uses this array directly. */
#if defined(__HIPCC__)
static __constant__ const double plom_bm_tab[PLOM_BM_TAB_SIZE] = PLOM_BM_TAB_VALUES;
#else
static const double plom_bm_tab[PLOM_BM_TAB_SIZE] = PLOM_BM_TAB_VALUES;
#endif

/* The coefficient table of the stream's Gaussian transform (gen_icdf_table.py); kernels
 * copy it to LDS and hand the stream a pointer to that copy. */
#if defined(__HIPCC__)
static __constant__ const double plom_icdf_tab[PLOM_ICDF_TAB_SIZE] = PLOM_ICDF_TAB_VALUES;
#else
static const double plom_icdf_tab[PLOM_ICDF_TAB_SIZE] = PLOM_ICDF_TAB_VALUES;
#endif

/* A particle's stream over one interval: its xoshiro128** state, the inverse-CDF table
 * (plom_icdf_tab or a copy of it) and the tables of the psr samplers: the log table (the
 * Box-Muller table plom_bm_tab or a copy of it) and log k! for k < 30 (plom_lfact_tab of
 * plom_psr.h or a copy of it); plom_nstream64_init sets the first and leaves the second
 * null, which plom_psr.h's users set. */
typedef struct { plom_xoshiro128 r; const double *tab; const double *ltab; const double *lftab; } plom_nstream64;

PLOM_HD void plom_nstream64_init(plom_nstream64 *ns, uint32_t seed, uint32_t stream, uint32_t n, uint32_t j,
                                 const double *tab)
{
    uint32_t w[4];
    plom_draw128(seed, stream, n, j, PLOM_PURPOSE_PROPAGATE, w);
    plom_xoshiro_seed(&ns->r, w);
    ns->tab = tab;
    ns->ltab = plom_bm_tab;
    ns->lftab = 0;
}

/* log x for a positive, normal, finite x by the log table: x = 2^e m, table entry k from
 * the top PLOM_BM_LOG_BITS fraction bits of m, log x = e ln2 + L_k + log1p(m invc_k - 1)
 * with the degree-7 Taylor polynomial of log1p on |z| < 2^-8 (truncation < 7e-21).
 * Entries k >= 64 hold L_k - ln2 and count one more power of two, so that log x near
 * x = 1 carries no cancellation.  No division and no branch.  The Box-Muller radius and
 * the psr samplers' logarithm. */
PLOM_HD double plom_log_tab(double x, const double *tab)
{
    const uint64_t b = plom_d2u(x);
    const uint32_t hi = (uint32_t)(b >> 32);
    const uint32_t k = (hi >> (20 - PLOM_BM_LOG_BITS)) & ((1u << PLOM_BM_LOG_BITS) - 1u);
    const int e = (int)(hi >> 20) - 1023 + (int)(k >> (PLOM_BM_LOG_BITS - 1));
    const double m = plom_u2d((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    const double z = plom_fma(m, tab[2 * k], -1.0);
    double q = plom_fma(z, 0x1.2492492492492p-3, -0x1.5555555555555p-3);
    q = plom_fma(q, z, 0x1.999999999999ap-3);
    q = plom_fma(q, z, -0x1.0000000000000p-2);
    q = plom_fma(q, z, 0x1.5555555555555p-2);
    q = plom_fma(q, z, -0x1.0000000000000p-1);
    const double p = plom_fma(z * z, q, z);
    const double de = (double)e;
    return plom_fma(de, PLOM_LN2_HI, tab[2 * k + 1]) + plom_fma(de, PLOM_LN2_LO, p);
}

/* -2 log u for u = (x + 1/2) 2^-32 */
PLOM_HD double plom_bm_m2log(uint32_t x, const double *tab)
{
    return -2.0 * plom_log_tab(plom_fma((double)x, 0x1p-32, 0x1p-33) /* exact */, tab);
}

/* sin and cos of 2 pi u for u = (y + 1/2) 2^-32: the angle is a_k + d with a_k =
 * 2 pi (k + 1/2)/256 (k = top 8 bits of y, tabulated) and d = 2 pi (r + 1/2) 2^-32,
 * r = (y mod 2^24) - 2^23, |d| <= pi/256; Taylor polynomials of sin d (to d^7) and
 * cos d - 1 (to d^6), truncation < 2e-20. */
#define PLOM_BM_DANG 0x1.921fb54442d18p-30 /* 2 pi 2^-32 */
PLOM_HD void plom_bm_sincos(uint32_t y, const double *tab, double *s, double *c)
{
    const uint32_t k = y >> (32 - PLOM_BM_SC_BITS);
    const int r = (int)(y & ((1u << (32 - PLOM_BM_SC_BITS)) - 1u)) - (1 << (31 - PLOM_BM_SC_BITS));
    const double d = plom_fma((double)r, PLOM_BM_DANG, 0.5 * PLOM_BM_DANG); /* (r + 1/2) 2 pi 2^-32 */
    const double sa = tab[PLOM_BM_SC_OFFSET + 2 * k], ca = tab[PLOM_BM_SC_OFFSET + 2 * k + 1];
    const double d2 = d * d;
    const double ps = plom_fma(plom_fma(d2, -0x1.a01a01a01a01ap-13, 0x1.1111111111111p-7), d2, -0x1.5555555555555p-3);
    const double sd = plom_fma(d * d2, ps, d);
    const double pc = plom_fma(plom_fma(d2, -0x1.6c16c16c16c17p-10, 0x1.5555555555555p-5), d2, -0x1.0000000000000p-1);
    const double cm1 = d2 * pc;
    *s = sa + plom_fma(ca, sd, sa * cm1);
    *c = ca + plom_fma(-sa, sd, ca * cm1);
}

/* one Box-Muller pair from the words x, y: z0 = r cos(2 pi u2), z1 = r sin(2 pi u2),
 * r = sqrt(-2 log u1) */
PLOM_HD void plom_bm_pair(uint32_t x, uint32_t y, const double *tab, double *z0, double *z1)
{
    const double rad = plom_sqrt_mid(plom_bm_m2log(x, tab));
    double s, c;
    plom_bm_sincos(y, tab, &s, &c);
    *z0 = rad * c;
    *z1 = rad * s;
}

/* The Gaussian of one 32-bit word w: v = w >> 1, u = (v + 1/2) 2^-32 in [2^-33, 1/2)
 * (exact), z = Phi^-1(u) from the polynomial of u's segment (binary octave of u and the
 * top PLOM_ICDF_SUB_BITS fraction bits: the top bits of the double's high word) at the
 * local coordinate t = 32 (1 + r) - 33 in [-1, 1) (r the remaining fraction, t exact),
 * and the sign flipped when w is odd.  About 2e-14 from Phi^-1(u) (checked against mpmath
 * by the tests); the mirror image of the low bit makes z symmetric exactly.  The table
 * holds segment s in slot (s + PLOM_ICDF_ROT64) mod NSEG, the low bits of the converted
 * word's exponent and top fraction bits, so the slot's byte offset is a shift and a mask
 * of the high word. */
#if PLOM_ICDF_DEG != 6
#error "plom_icdf_normal evaluates a degree-6 table"
#endif
#define PLOM_ICDF_T_SCALE ((double)(2 << PLOM_ICDF_SUB_BITS))
PLOM_HD double plom_icdf_normal(uint32_t w, const double *tab)
{
    /* (v + 1/2) 2^-32 = (2v + 1) 2^-33 = (w | 1) 2^-33 = u.  The double v1 = w | 1 (exact)
     * has u's fraction bits and a biased exponent 33 higher, so the segment is read from v1
     * directly and u itself is never formed. */
    const double v1 = (double)(w | 1u);
    const uint64_t b = plom_d2u(v1);
    const uint32_t hi = (uint32_t)(b >> 32);
    /* the slot of u's segment (E + 33) 16 + s: v1's biased exponent E + 33 + 1023 and top
     * fraction bits modulo the table size, as a byte offset (the mask also lets the
     * compiler fold the coefficients' offsets into the LDS loads) */
    const uint32_t off = (hi >> (20 - PLOM_ICDF_SUB_BITS - 3)) & ((PLOM_ICDF_NSEG - 1u) << 3);
    /* x = 32 m, m in [1, 1 + 2^-4) the remaining fraction r after the top fraction bits
     * (1 + r with v1's exponent replaced by that of 32: the same bits, exact), and
     * t = 32 (1 + r) - 33 = x - 33, exact: x has at most 48 fraction bits below 2^5, so
     * x - 33 in [-1, 1) needs 48 bits */
    const uint32_t keep = (1u << (20 - PLOM_ICDF_SUB_BITS)) - 1u;
    const double x = plom_u2d(((uint64_t)((hi & keep) | ((1023u + PLOM_ICDF_SUB_BITS + 1u) << 20)) << 32) | (b & 0xffffffffULL));
    const double t = x - (PLOM_ICDF_T_SCALE + 1.0);
    const double *c = (const double *)((const char *)tab + off);
    double p = c[6 * PLOM_ICDF_NSEG];
    p = plom_fma(p, t, c[5 * PLOM_ICDF_NSEG]);
    p = plom_fma(p, t, c[4 * PLOM_ICDF_NSEG]);
    p = plom_fma(p, t, c[3 * PLOM_ICDF_NSEG]);
    p = plom_fma(p, t, c[2 * PLOM_ICDF_NSEG]);
    p = plom_fma(p, t, c[1 * PLOM_ICDF_NSEG]);
    p = plom_fma(p, t, c[0]);
    return plom_u2d(plom_d2u(p) ^ ((uint64_t)(w & 1u) << 63));
}

/* the next two Gaussians of the stream, in draw order */
PLOM_HD void plom_normal_pair64(plom_nstream64 *ns, double *z0, double *z1)
{
    const uint32_t x = plom_xoshiro_next(&ns->r);
    const uint32_t y = plom_xoshiro_next(&ns->r);
    *z0 = plom_icdf_normal(x, ns->tab);
    *z1 = plom_icdf_normal(y, ns->tab);
}

/* the uniform that picks the final particle of a lineage, as the numerator 2x+1 of
 * U = (2x+1)/2^33 */
PLOM_HD uint64_t plom_lineage_num(uint32_t seed, uint32_t stream)
{
    uint32_t w[4];
    plom_draw128(seed, stream, 0u, 0u, PLOM_PURPOSE_FULL, w);
    return 2ull * (uint64_t)w[0] + 1ull;
}

/* one observation row n of slot s: out[2 ts] = mean, out[2 ts + 1] = max(mean + sd z, 0) with
 * z the Box-Muller values of Philox(ctr = (s, n, 3, 0), key = (seed, stream)), two series per
 * 128-bit draw (ctr word 3 counts the draws); slot 0 is the lineage, slot j + 1 particle j of
 * the stored clouds */
PLOM_HD void plom_obs_draw(uint32_t seed, uint32_t stream, uint32_t n, uint32_t slot, int n_ts, const double *mean,
                           const double *var, const double *tab, double *out)
{
    for (int ts = 0; ts < n_ts; ts += 2) {
        uint32_t w[4];
        const uint32_t ctr[4] = { slot, n, PLOM_PURPOSE_OBS, (uint32_t)(ts / 2) };
        const uint32_t key[2] = { seed, stream };
        plom_philox4x32_10(ctr, key, w);
        double z[2];
        plom_bm_pair(w[0], w[1], tab, &z[0], &z[1]);
        for (int k = 0; k < 2 && ts + k < n_ts; k++) {
            const double y = mean[ts + k] + plom_sqrt(var[ts + k]) * z[k];
            out[2 * (ts + k)] = mean[ts + k];
            out[2 * (ts + k) + 1] = (y > 0.0) ? y : 0.0;
        }
    }
}

/* the observation row n of a lineage (slot 0) */
PLOM_HD void plom_lineage_obs(uint32_t seed, uint32_t stream, uint32_t n, int n_ts, const double *mean,
                              const double *var, const double *tab, double *out)
{
    plom_obs_draw(seed, stream, n, 0u, n_ts, mean, var, tab, out);
}

/* the resampling uniform of interval n as the integer numerator 2x+1 of U = (2x+1)/2^33 */
PLOM_HD uint64_t plom_resample_num(uint32_t seed, uint32_t stream, uint32_t n)
{
    uint32_t w[4];
    plom_draw128(seed, stream, n, 0u, PLOM_PURPOSE_RESAMPLE, w);
    return 2ull * (uint64_t)w[0] + 1ull;
}

/* ---------------------------------------------------------------- FP32 Gaussians */
#if defined(__HIPCC__)
/* A particle's FP32 stream: the same xoshiro128** words as the FP64 stream and the same
 * inverse-CDF transform in FP32 (the table's coefficients rounded to float), so that an
 * FP32 filter draws the FP64 filter's Gaussians rounded to FP32. */
typedef struct { plom_xoshiro128 r; const float *tab; } plom_nstream32;

/* the coefficients rounded to float; kernels copy them to LDS */
static __constant__ const float plom_icdf_tab32[PLOM_ICDF_TAB_SIZE] = PLOM_ICDF_TAB32_VALUES;

__device__ inline __attribute__((always_inline)) void
plom_nstream32_init(plom_nstream32 *ns, uint32_t seed, uint32_t stream, uint32_t n, uint32_t j, const float *tab)
{
    uint32_t w[4];
    plom_draw128(seed, stream, n, j, PLOM_PURPOSE_PROPAGATE, w);
    plom_xoshiro_seed(&ns->r, w);
    ns->tab = tab;
}

/* plom_icdf_normal in FP32 on the top 24 bits of the word: u = ((w >> 9) + 1/2) 2^-24 =
 * ((w >> 8) | 1) 2^-25 (exact in float, within 2^-25 of the FP64 transform's u, in
 * [2^-25, 1/2)), the same segments (the float's exponent and top 4 fraction bits; the FP32
 * table holds segment s in slot (s + PLOM_ICDF_ROT32) mod NSEG) and the same sign bit
 * w & 1 */
__device__ inline __attribute__((always_inline)) float plom_icdf_normal32(uint32_t w, const float *tab)
{
    /* the float v1 = (w >> 8) | 1 (exact) has u's fraction bits and a biased exponent 25
     * higher: the segment is read from v1 directly */
    const float v1 = (float)((w >> 8) | 1u);
    const uint32_t b = __float_as_uint(v1);
    const uint32_t off = (b >> (23 - PLOM_ICDF_SUB_BITS - 2)) & ((PLOM_ICDF_NSEG - 1u) << 2);
    const float m = __uint_as_float((b & ((1u << (23 - PLOM_ICDF_SUB_BITS)) - 1u)) | 0x3f800000u);
    /* t = 32 (m - 1) - 1 = 32 m - 33 in one fused multiply-add, exact: m in [1, 1 + 2^-4)
     * has 19 fraction bits, so 32 m - 33 in [-1, 1) needs 19 bits */
    const float t = __builtin_fmaf(m, (float)PLOM_ICDF_T_SCALE, -(float)(PLOM_ICDF_T_SCALE + 1.0));
    const float *c = (const float *)((const char *)tab + off);
    float p = c[6 * PLOM_ICDF_NSEG];
    p = __builtin_fmaf(p, t, c[5 * PLOM_ICDF_NSEG]);
    p = __builtin_fmaf(p, t, c[4 * PLOM_ICDF_NSEG]);
    p = __builtin_fmaf(p, t, c[3 * PLOM_ICDF_NSEG]);
    p = __builtin_fmaf(p, t, c[2 * PLOM_ICDF_NSEG]);
    p = __builtin_fmaf(p, t, c[1 * PLOM_ICDF_NSEG]);
    p = __builtin_fmaf(p, t, c[0]);
    return __uint_as_float(__float_as_uint(p) ^ ((w & 1u) << 31));
}

__device__ inline __attribute__((always_inline)) void plom_normal_pair32(plom_nstream32 *ns, float *z0, float *z1)
{
    const uint32_t x = plom_xoshiro_next(&ns->r);
    const uint32_t y = plom_xoshiro_next(&ns->r);
    *z0 = plom_icdf_normal32(x, ns->tab);
    *z1 = plom_icdf_normal32(y, ns->tab);
}
#endif

#endif /* PLOM_RNG_H */
