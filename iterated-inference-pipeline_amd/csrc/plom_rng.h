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
 * Gaussians: Box-Muller pairs on two uniforms; the generated step code takes the two
 * values of a pair in draw order and drops the second value of a pair left open at the
 * end of an Euler substep, so the pairing is fixed at code-generation time.  FP64 works
 * on the two 32-bit words with small lookup tables (plom_bm_m2log, plom_bm_sincos: no
 * division, no branch) and IEEE sqrt; the FP32 stream (device only) uses the gfx950
 * transcendental instructions.
 */
#ifndef PLOM_RNG_H
#define PLOM_RNG_H

#include "plom_math.h"
#include "plom_bm_tables.h"

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

typedef struct { plom_xoshiro128 r; const double *tab; } plom_nstream64;

PLOM_HD void plom_nstream64_init(plom_nstream64 *ns, uint32_t seed, uint32_t stream, uint32_t n, uint32_t j,
                                 const double *tab)
{
    uint32_t w[4];
    plom_draw128(seed, stream, n, j, PLOM_PURPOSE_PROPAGATE, w);
    plom_xoshiro_seed(&ns->r, w);
    ns->tab = tab;
}

/* -2 log u for u = (x + 1/2) 2^-32: u = 2^e m, table entry k from the top fraction bits
 * of m, log u = e ln2 + L_k + log1p(m invc_k - 1) with the degree-7 Taylor polynomial of
 * log1p on |z| < 2^-8 (truncation < 7e-21).  Entries k >= 64 hold L_k - ln2 and count
 * one more power of two, so that log u near u = 1 carries no cancellation. */
PLOM_HD double plom_bm_m2log(uint32_t x, const double *tab)
{
    const double u = plom_fma((double)x, 0x1p-32, 0x1p-33); /* exact */
    const uint64_t b = plom_d2u(u);
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
    return -2.0 * (plom_fma(de, PLOM_LN2_HI, tab[2 * k + 1]) + plom_fma(de, PLOM_LN2_LO, p));
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

PLOM_HD void plom_normal_pair64(plom_nstream64 *ns, double *z0, double *z1)
{
    const uint32_t x = plom_xoshiro_next(&ns->r);
    const uint32_t y = plom_xoshiro_next(&ns->r);
    plom_bm_pair(x, y, ns->tab, z0, z1);
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
typedef plom_xoshiro128 plom_nstream32;

__device__ inline __attribute__((always_inline)) void
plom_nstream32_init(plom_nstream32 *ns, uint32_t seed, uint32_t stream, uint32_t n, uint32_t j)
{
    uint32_t w[4];
    plom_draw128(seed, stream, n, j, PLOM_PURPOSE_PROPAGATE, w);
    plom_xoshiro_seed(ns, w);
}

/* u in (0,1) from the top 24 bits: exact in float */
__device__ inline __attribute__((always_inline)) float plom_u01f(uint32_t x)
{
    return ((float)(x >> 8) + 0.5f) * 0x1p-24f;
}

__device__ inline __attribute__((always_inline)) void plom_normal_pair32(plom_nstream32 *ns, float *z0, float *z1)
{
    float u1 = plom_u01f(plom_xoshiro_next(ns));
    float u2 = plom_u01f(plom_xoshiro_next(ns));
    /* -2 ln u1 = -2 ln2 log2(u1); v_log_f32 is log2, v_sin/v_cos_f32 take turns */
    float rad = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
    *z0 = rad * __builtin_amdgcn_cosf(u2);
    *z1 = rad * __builtin_amdgcn_sinf(u2);
}
#endif

#endif /* PLOM_RNG_H */
