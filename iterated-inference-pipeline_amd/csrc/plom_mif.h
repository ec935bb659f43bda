/*
 * plom_mif.h -- arithmetic of iterated filtering shared by the GPU kernels and the CPU
 * oracle (compiled with -ffp-contract=off on both sides, so the results are bit-identical).
 *
 *   plom_mif_natural   back_transform_x (C/core/transform.c:462-480): the inverse transform
 *                      (f_inv_log, f_inv_logit, f_inv_logit_ab, transform.c:40-100), the
 *                      scale 10^x or 10^-x, the unit multiplier, and 1/(1e-5 + x) for a
 *                      duration; with this library's exp (plom_math.h).
 *   plom_mif_normals   the two standard normals of the random walk of components
 *                      2b and 2b+1 of particle j after interval n-1 (the draw of
 *                      resample_and_mut_theta_mif, C/mif/methods.c:339-346): Box-Muller of
 *                      Philox-4x32-10(ctr = (j, n, 4, b), key = (seed, stream)).
 *   plom_mif_combine   count, mean and sum of squared deviations of the union of two
 *                      disjoint sets of values (Chan, Golub and LeVeque 1983): the merge
 *                      step of the fixed-shape tree over tiles that gives the variance of
 *                      theta (the reference's one-pass variance, methods.c:168-181).
 *   plom_mif_prior_pdf the prior densities on the natural scale: uniform on [a, b)
 *                      (gsl_ran_flat_pdf), normal with mean (a+b)/2 and sd (b-a)/4
 *                      (normal_prior, C/core/metropolis_hastings_prior.c:93-98) and the
 *                      pseudo-uniform density with quadratic shoulders of width
 *                      d = (b-a)/100 (pseudo_unif_prior, same file:109-132).
 *   plom_mif_prior_term  e log(max(p, LIKE_MIN)) for one component (NaN, infinite or
 *                      negative densities count as LIKE_MIN, sanitize_likelihood,
 *                      C/core/likelihood.c:77-91): the log of the factor prior^e by which
 *                      mif --prior multiplies a particle's weight (patch_likelihood_prior,
 *                      C/mif/methods.c:469-523).
 *   plom_mif_prior_patch  the patched log-weight, or -inf for a failure (weight(),
 *                      C/core/smc.c:44: a weight at or below LIKE_MIN^N_TS).
 */
#ifndef PLOM_MIF_H
#define PLOM_MIF_H

#include "plom_math.h"
#include "plom_rng.h"

#define PLOM_PURPOSE_MUTATE 4u
#define PLOM_MIF_MVN_MAX 64     /* components of a correlated random walk */
#define PLOM_MIF_LN10 2.302585092994046

PLOM_HD double plom_mif_f_inv(int tr, double x, double a, double b)
{
    double e;
    switch (tr) {
    case 1:
        return plom_exp(x);
    case 2:
        if (x > 0.0) return 1.0 / (1.0 + plom_exp(-x));
        e = plom_exp(x);
        return e / (1.0 + e);
    case 3:
        if (a == b) return x;
        if (x < 0.0) {
            e = plom_exp(x);
            return (b * e + a) / (1.0 + e);
        }
        e = plom_exp(-x);
        return (b + a * e) / (1.0 + e);
    default:
        return x;
    }
}

PLOM_HD double plom_mif_natural(int tr, int sc, int dur, double x, double a, double b, double mult)
{
    double y = plom_mif_f_inv(tr, x, a, b);
    if (sc == 1) y = plom_exp(y * PLOM_MIF_LN10);
    else if (sc == 2) y = plom_exp(-y * PLOM_MIF_LN10);
    y = y * mult;
    if (dur) y = 1.0 / (0.00001 + y);
    return y;
}

PLOM_HD void plom_mif_normals(uint32_t seed, uint32_t stream, uint32_t n, uint32_t j, uint32_t b, const double *tab,
                              double *z0, double *z1)
{
    uint32_t w[4];
    const uint32_t ctr[4] = { j, n, PLOM_PURPOSE_MUTATE, b };
    const uint32_t key[2] = { seed, stream };
    plom_philox4x32_10(ctr, key, w);
    plom_bm_pair(w[0], w[1], tab, z0, z1);
}

#ifndef PLOM_PRIOR_NONE          /* (also in include/plom_smc.h) */
#define PLOM_PRIOR_NONE (-1)
#define PLOM_PRIOR_UNIFORM 0
#define PLOM_PRIOR_NORMAL 1
#define PLOM_PRIOR_PSEUDO_UNIFORM 2
#endif
#define PLOM_MIF_INV_SQRT_2PI 0.3989422804014327

PLOM_HD double plom_mif_prior_pdf(int kind, double x, double a, double b)
{
    if (kind == PLOM_PRIOR_NORMAL) {
        const double sd = (b - a) / 4.0;
        const double u = (x - 0.5 * (a + b)) / sd;
        return PLOM_MIF_INV_SQRT_2PI / sd * plom_exp(-0.5 * (u * u));
    }
    if (kind == PLOM_PRIOR_PSEUDO_UNIFORM) {
        /* flat at h on [a+d, b-d]; on the shoulders the parabola c2 (x-a-d)^2 + h, which is 0
         * at a and b (mirrored on the right) and has the flat part's value and slope at the
         * joins; h makes the density integrate to 1 */
        if (x <= a || x >= b) return 0.0;
        const double d = 0.01 * (b - a);
        const double h = 1.0 / ((b - a) - (2.0 / 3.0) * d);
        double u;
        if (x < a + d) u = x - (a + d);
        else if (x < b - d) return h;
        else u = x - (b - d);
        return h - (h / (d * d)) * (u * u);
    }
    return (x >= a && x < b) ? 1.0 / (b - a) : 0.0;
}

PLOM_HD double plom_mif_prior_term(int kind, double e, int tr, double theta, double a, double b, double like_min)
{
    double p = plom_mif_prior_pdf(kind, plom_mif_f_inv(tr, theta, a, b), a, b);
    if (!(p > like_min) || p == PLOM_INF) p = like_min;
    return e * plom_log(p);
}

/* l: the particle's log-likelihood (the floored sum when every series is observed and
 * floored: all_floored); pt: its summed prior terms; all_fail_llike: N_TS log LIKE_MIN */
PLOM_HD double plom_mif_prior_patch(double l, int all_floored, double pt, double all_fail_llike)
{
    if (all_floored) return pt > 0.0 ? l + pt : -PLOM_INF;
    const double lp = l + pt;
    return lp <= all_fail_llike ? -PLOM_INF : lp;
}

/* (c, mean, m2) of a followed by b; an empty side (c == 0) leaves the other unchanged */
PLOM_HD void plom_mif_combine(double ca, double ma, double qa, double cb, double mb, double qb, double *c, double *m,
                              double *q)
{
    if (ca == 0.0) { *c = cb; *m = mb; *q = qb; return; }
    if (cb == 0.0) { *c = ca; *m = ma; *q = qa; return; }
    const double n = ca + cb;
    const double d = mb - ma;
    *c = n;
    *m = ma + d * (cb / n);
    *q = (qa + qb) + (d * d) * ((ca * cb) / n);
}

#endif
