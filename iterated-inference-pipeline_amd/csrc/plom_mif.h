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
