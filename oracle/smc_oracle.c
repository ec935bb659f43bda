/*
 * smc_oracle.c -- CPU restatement of the PLOM SMC filter: TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker of the HIP filter, never the product: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * It restates, as plain sequential C:
 *   run_SMC_zmq_inproc      model_builder/C/core/smc.c:379-467 (driver order: propagate
 *                           every particle, weight, select, gather)
 *   worker j-loop           model_builder/C/core/worker_inproc.c:71-80 (reset_inc ->
 *                           f_pred -> proj2obs -> weight)
 *   weight                  model_builder/C/core/smc.c:27-83
 *   systematic_sampling     model_builder/C/core/smc.c:90-112
 *   resample_X              model_builder/C/core/smc.c:142-161
 *   one MIF pass            model_builder/C/mif/mif.c:128-200 with per-particle parameters
 *                           (oracle_mif_pass): the particles' theta resampled with their
 *                           state and moved by the random walk after every observation
 *                           (resample_and_mut_theta_mif, C/mif/methods.c:261-367), and the
 *                           weighted mean / variance of theta per observation
 *                           (mean_var_theta_theoretical_mif, methods.c:146-205) in the
 *                           canonical tree order of the GPU (tiles of 256, then tiles);
 *                           with a prior, the weights patched by prior^e per component
 *                           (patch_likelihood_prior, methods.c:469-523) in the log domain
 * Two arithmetic modes:
 *   ORACLE_CANONICAL (0)    the arithmetic the GPU filter defines (log-domain weights,
 *                           recursive pairwise tree sums over tiles of 256, fixed-point
 *                           2^62 weights and an exact integer systematic resampling), each
 *                           step written here independently of the kernels: sequential
 *                           recursion for the trees, a running integer prefix and the
 *                           reference's two-pointer walk for the selection;
 *   ORACLE_REFERENCE (1)    the reference's own arithmetic: raw likelihoods exp(l), the
 *                           failure test w <= LIKE_MIN^N_TS, sequential double sums,
 *                           log(sum/J), weight_cum += w_i and ran += 1/J (smc.c:44-110),
 *                           with the index clamped at J-1 (smc.c:105 has no bound check).
 * Propagation and observation use the generated model code and the shared math/RNG
 * headers (csrc/plom_math.h, csrc/plom_rng.h), which are pinned separately (Philox
 * known-answer vectors, libm/mpmath accuracy, closed-form Euler tests).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "plom_math.h"
#include "plom_rng.h"
#include "plom_mif.h"
#include "plom_psr.h"
#include "plom_cov.h"
#include "plom_models.h"

#define ORACLE_CANONICAL 0
#define ORACLE_REFERENCE 1
#define TILE 256
#define FLAG_NO_FILTER 1u

typedef unsigned __int128 u128;

typedef void (*interval_fn)(double *, const double *, plom_nstream64 *, double, double, double, double, unsigned int,
                            const double *, int, double, double, const plom_drift_view *, const plom_spline_set *);
typedef void (*proj2obs_fn)(const double *, const double *, const double *, double *);
typedef double (*loglik_fn)(const double *, const double *, const double *, const double *, double, double, int *);
typedef void (*moments_fn)(const double *, const double *, const double *, double, double *, double *);
typedef void (*hat_fn)(const double *, const double *, const double *, const plom_drift_view *, double, double *);

typedef struct {
    const char *name;
    int S, n_ts, n_hat, n_pp, pv_size, fix_off, n_fix, n_drift, n_cac;
    interval_fn interval;
    proj2obs_fn proj2obs;
    loglik_fn loglik;
    moments_fn obs_moments;
    hat_fn hat_values;
} oracle_model;

#define ORACLE_MODEL(lname, UNAME)                                                                 \
    { #lname, UNAME##_SIZE_PROJ, UNAME##_N_TS, UNAME##_N_HAT, UNAME##_N_PP, UNAME##_PV_SIZE,           \
      UNAME##_FIX_OFF, UNAME##_N_FIX, UNAME##_N_DRIFT, UNAME##_N_CAC,                                  \
      plom_##lname##_interval_f64,                                                                   \
      plom_##lname##_proj2obs_f64, plom_##lname##_loglik, plom_##lname##_obs_moments, plom_##lname##_hat_values }

/* an extra model (a user's model outside the library, built from its generated header in
 * another directory): -DORACLE_EXTRA_MODEL=<name> -DORACLE_EXTRA_UNAME=PLOM_<NAME>
 * -DORACLE_EXTRA_HEADER='"plom_<name>.h"'; it takes the id after the built-in models */
#define ORACLE_MODEL_X(lname, UNAME) ORACLE_MODEL(lname, UNAME)
#ifdef ORACLE_EXTRA_MODEL
#include ORACLE_EXTRA_HEADER
#endif

static const oracle_model MODELS[] = {
    ORACLE_MODEL(sir, PLOM_SIR),
    ORACLE_MODEL(sirs4, PLOM_SIRS4),
    ORACLE_MODEL(seir10, PLOM_SEIR10),
    ORACLE_MODEL(noise, PLOM_NOISE),
    ORACLE_MODEL(drift, PLOM_DRIFT),
#ifdef ORACLE_EXTRA_MODEL
    ORACLE_MODEL_X(ORACLE_EXTRA_MODEL, ORACLE_EXTRA_UNAME),
#endif
};
#define N_MODELS ((int)(sizeof MODELS / sizeof MODELS[0]))

int oracle_model_count(void) { return N_MODELS; }
const char *oracle_model_name(int m) { return (m >= 0 && m < N_MODELS) ? MODELS[m].name : ""; }
int oracle_model_size_proj(int m) { return MODELS[m].S; }
int oracle_model_n_ts(int m) { return MODELS[m].n_ts; }
int oracle_model_n_hat(int m) { return MODELS[m].n_hat; }
int oracle_model_n_pp(int m) { return MODELS[m].n_pp; }
/* count draws of one sampler from the particle stream (seed, stream, n, j), in order:
 * kind 0 uniform (two words), 1 Gamma(a, 1), 2 Poisson(a), 3 Binomial(a, b), 4 log a!, 6 uniform
 * (one word), 7 log a! + log b! (plom_lfact2); for kind 5 the multinomial of N = a over the K = count probabilities in p,
 * written to out[0..K) */
void oracle_psr_draws(int kind, double a, double b, uint32_t seed, uint32_t stream, uint32_t n, uint32_t j, int count,
                      const double *p, double *out)
{
    plom_nstream64 ns;
    plom_nstream64_init(&ns, seed, stream, n, j, plom_icdf_tab);
    plom_psr_tables(&ns, plom_bm_tab, plom_lfact_tab);
    if (kind == 5) {
        plom_multinomial(&ns, count, a, p, out);
        return;
    }
    for (int i = 0; i < count; i++) {
        switch (kind) {
        case 0: out[i] = plom_unif53(&ns); break;
        case 1: out[i] = plom_gamma_draw(&ns, a); break;
        case 2: out[i] = plom_poisson_draw(&ns, a); break;
        case 3: out[i] = plom_binom_draw(&ns, a, b); break;
        case 6: out[i] = plom_unif32(&ns); break;
        case 7: out[i] = plom_lfact2(a + (double)i, b, plom_bm_tab, plom_lfact_tab); break;
        default: out[i] = plom_lfact(a + (double)i, plom_bm_tab, plom_lfact_tab); break;
        }
    }
}

double oracle_mif_natural(int tr, int sc, int dur, double x, double a, double b, double mult)
{
    return plom_mif_natural(tr, sc, dur, x, a, b, mult);
}

double oracle_mif_prior_pdf(int kind, double x, double a, double b) { return plom_mif_prior_pdf(kind, x, a, b); }

double oracle_mif_prior_term(int kind, double e, int tr, double theta, double a, double b, double like_min)
{
    return plom_mif_prior_term(kind, e, tr, theta, a, b, like_min);
}

/* ------------------------------------------------------------------ covariates */
/* The covariate table of a filter (plom_cov.h): the rows of interval n's substeps and the
 * row of observation time times[r]; without time-varying covariates, the covariate slots of
 * pv repeated (cs = 0). */
typedef struct {
    const double *tab;
    int cs, kmax, N;
    double *owned;
    /* the splines themselves (plom_spline_pack), for the ode implementation */
    plom_spline_set spl;
    int32_t *spl_n, *spl_off;
    double *spl_val;
} cov_view;

static void cov_free(cov_view *c)
{
    free(c->owned);
    free(c->spl_n);
    free(c->spl_off);
    free(c->spl_val);
}

/* the splines for an ode run, NULL without time-varying covariates */
static const plom_spline_set *cv_splines(const cov_view *c, uint32_t noises_off)
{
    return (c->spl_val && (noises_off & PLOM_IMPL_ODE)) ? &c->spl : NULL;
}

static const double *cv_steps(const cov_view *c, int n) { return c->tab + (size_t)n * (size_t)c->kmax * (size_t)c->cs; }
static const double *cv_at(const cov_view *c, int r)
{
    return c->tab + ((size_t)c->N * (size_t)c->kmax + (size_t)r) * (size_t)c->cs;
}

/* cov_n [n_fix] knot counts, cov_x / cov_y the knots of every series one after another;
 * cov_n NULL: the constants of pv.  0, or -3 when a time falls outside a series' knots. */
static int cov_build(const oracle_model *M, int N, const double *times, double dt, const int32_t *cov_n,
                     const double *cov_x, const double *cov_y, const double *pv, cov_view *out)
{
    memset(out, 0, sizeof *out);
    out->N = N;
    if (!cov_n || M->n_fix == 0) {
        out->tab = pv + M->fix_off;
        return 0;
    }
    plom_covariate *cv = malloc(sizeof(plom_covariate) * (size_t)M->n_fix);
    if (!cv) return -2;
    size_t o = 0;
    for (int i = 0; i < M->n_fix; i++) {
        cv[i].n = cov_n[i];
        cv[i].pad = 0;
        cv[i].x = cov_x + o;
        cv[i].y = cov_y + o;
        o += (size_t)cov_n[i];
    }
    out->kmax = plom_cov_kmax(N, times, dt);
    out->cs = M->n_fix;
    out->owned = malloc(sizeof(double) * plom_cov_rows(N, out->kmax) * (size_t)M->n_fix);
    char msg[256];
    int rc = out->owned ? plom_cov_table(cv, M->n_fix, N, times, dt, out->kmax, out->owned, msg, sizeof msg) : -2;
    if (rc == 0) {
        out->spl_n = malloc(sizeof(int32_t) * (size_t)M->n_fix);
        out->spl_off = malloc(sizeof(int32_t) * (size_t)M->n_fix);
        out->spl_val = malloc(sizeof(double) * (plom_spline_vals(cv, M->n_fix) + 1));
        rc = (out->spl_n && out->spl_off && out->spl_val)
                 ? plom_spline_pack(cv, M->n_fix, out->spl_n, out->spl_off, out->spl_val, msg, sizeof msg) : -2;
        out->spl.n = out->spl_n;
        out->spl.off = out->spl_off;
        out->spl.val = out->spl_val;
    }
    free(cv);
    out->tab = out->owned;
    return rc == 0 ? 0 : -3;
}

/* the table itself (for tests): out[plom_cov_rows(N, kmax) * n_fix], *kmax_out */
int oracle_cov_table(int model, int N, const double *times, double dt, const int32_t *cov_n, const double *cov_x,
                     const double *cov_y, double *out, int32_t *kmax_out)
{
    if (model < 0 || model >= N_MODELS) return -1;
    cov_view c;
    const int rc = cov_build(&MODELS[model], N, times, dt, cov_n, cov_x, cov_y, NULL, &c);
    if (rc) { cov_free(&c); return rc; }
    *kmax_out = c.kmax;
    memcpy(out, c.tab, sizeof(double) * plom_cov_rows(N, c.kmax) * (size_t)MODELS[model].n_fix);
    cov_free(&c);
    return 0;
}

/* one spline (plom_cov.h) at the times t[0..nt) (for tests) */
int oracle_spline(int n, const double *x, const double *y, int nt, const double *t, double *out)
{
    double *c = malloc(sizeof(double) * (size_t)(n + 1));
    if (!c || plom_spline_prepare(n, x, y, c)) { free(c); return -2; }
    for (int i = 0; i < nt; i++) out[i] = plom_spline_eval(n, x, y, c, t[i]);
    free(c);
    return 0;
}

int oracle_model_n_fix(int m) { return MODELS[m].n_fix; }

/* ------------------------------------------------------------------ drifting parameters */
/* the view of a drift map given as flat arrays [n_drift * n_cac] (group NULL: no drift) */
typedef struct {
    plom_drift_view v;
    int32_t ng[64];
    int on;
} drift_arg;

static void drift_build(const oracle_model *M, const int32_t *group, const int32_t *tr, const int32_t *sc,
                        const int32_t *dur, const double *mn, const double *mx, const double *mult, const int32_t *vol,
                        drift_arg *out)
{
    memset(out, 0, sizeof *out);
    if (!group || M->n_drift == 0) return;
    for (int i = 0; i < M->n_drift && i < 64; i++) {
        out->ng[i] = 0;
        for (int c = 0; c < M->n_cac; c++)
            if (group[i * M->n_cac + c] + 1 > out->ng[i]) out->ng[i] = group[i * M->n_cac + c] + 1;
    }
    out->v.group = group; out->v.ng = out->ng; out->v.tr = tr; out->v.sc = sc; out->v.dur = dur;
    out->v.min = mn; out->v.max = mx; out->v.mult = mult; out->v.vol_slot = vol;
    out->on = 1;
}

static const plom_drift_view *drift_view(const drift_arg *d) { return d->on ? &d->v : NULL; }
int oracle_model_fix_off(int m) { return MODELS[m].fix_off; }

/* ------------------------------------------------------------------ canonical trees */
/* recursive contiguous pairwise sum of a[0..n), n a power of two */
static double tree_sum(const double *a, int n)
{
    if (n == 1) return a[0];
    return tree_sum(a, n / 2) + tree_sum(a + n / 2, n / 2);
}

static int next_pow2(int n)
{
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

/* ------------------------------------------------------------------ propagation */
typedef struct {
    const oracle_model *M;
    double *X;          /* [J][S] */
    double *logw;       /* [J] */
    const double *pv, *times, *data;
    size_t pv_stride;   /* 0: one pv for every particle; else particle j uses pv + j * pv_stride */
    const cov_view *cv;
    const plom_drift_view *dm;
    int J, n, nonan, filter, n_ts;
    uint32_t seed, stream, noises_off;
    double dt, sqdt, like_min, eps_abs, eps_rel;
    int j0, j1, mode;
    const double *lprior;   /* [J] or NULL: summed prior terms of the particles (mif --prior) */
    double all_fail_llike;  /* n_ts log(like_min), with lprior */
} prop_job;

static void propagate_range(prop_job *p)
{
    const oracle_model *M = p->M;
    const int S = M->S;
    double obs[64];
    for (int j = p->j0; j < p->j1; j++) {
        double *x = p->X + (size_t)j * S;
        const double *pv = p->pv + (size_t)j * p->pv_stride;
        plom_nstream64 ns;
        plom_nstream64_init(&ns, p->seed, p->stream, (uint32_t)p->n, (uint32_t)j, plom_icdf_tab);
        plom_psr_tables(&ns, plom_bm_tab, plom_lfact_tab);
        M->interval(x, pv, &ns, p->times[p->n], p->times[p->n + 1], p->dt, p->sqdt, p->noises_off, cv_steps(p->cv, p->n),
                    p->cv->cs, p->eps_abs, p->eps_rel, p->dm, cv_splines(p->cv, p->noises_off));
        if (p->filter && p->nonan) {
            int nf = 0;
            const double *cvd = cv_at(p->cv, p->n + 1);
            M->proj2obs(x, pv, cvd, obs);
            double l = M->loglik(obs, p->data + (size_t)p->n * p->n_ts, pv, cvd, p->times[p->n + 1], p->like_min, &nf);
            if (p->lprior) {
                /* patch_likelihood_prior (C/mif/methods.c:469-523) in the log domain: the
                 * weight exp(l) prior^e fails at or below like_min^n_ts (smc.c:44) */
                const double pt = p->lprior[j];
                if (nf == p->n_ts) l = pt > 0.0 ? l + pt : -INFINITY;
                else l = (l + pt <= p->all_fail_llike) ? -INFINITY : l + pt;
            } else if (p->mode == ORACLE_CANONICAL && nf == p->n_ts) {
                l = -INFINITY;
            }
            p->logw[j] = l;
        }
    }
}

static void *prop_thread(void *arg)
{
    propagate_range((prop_job *)arg);
    return NULL;
}

static void propagate_all(prop_job *base, int nthreads)
{
    if (nthreads <= 1 || base->J < 2 * nthreads) {
        base->j0 = 0;
        base->j1 = base->J;
        propagate_range(base);
        return;
    }
    pthread_t th[256];
    prop_job jobs[256];
    if (nthreads > 256) nthreads = 256;
    int chunk = base->J / nthreads;
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = *base;
        jobs[t].j0 = t * chunk;
        jobs[t].j1 = (t + 1 == nthreads) ? base->J : (t + 1) * chunk;
        pthread_create(&th[t], NULL, prop_thread, &jobs[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------ weighting + selection */
/* canonical: returns 1 if resampled (select filled), 0 for the all-fail branch */
static int weight_select_canonical(const double *logw, int J, int n, uint32_t seed, uint32_t stream,
                                   double all_fail_llike, double *llike_n, double *ess_n, uint32_t *select,
                                   double *work, uint64_t *cum)
{
    const int T = (J + TILE - 1) / TILE;
    double *tmax = work, *tsum = work + T, *tsum2 = work + 2 * T, *buf = work + 3 * T;
    double m = -INFINITY;
    for (int t = 0; t < T; t++) {
        double mt = -INFINITY;
        for (int i = 0; i < TILE; i++) {
            int j = t * TILE + i;
            if (j < J && logw[j] > mt) mt = logw[j];
        }
        for (int i = 0; i < TILE; i++) {
            int j = t * TILE + i;
            double w = 0.0;
            if (j < J && mt != -INFINITY && logw[j] != -INFINITY) w = plom_exp(logw[j] - mt);
            buf[i] = w;
        }
        tsum[t] = tree_sum(buf, TILE);
        for (int i = 0; i < TILE; i++) buf[i] = buf[i] * buf[i];
        tsum2[t] = tree_sum(buf, TILE);
        tmax[t] = mt;
        if (mt > m) m = mt;
    }
    if (m == -INFINITY) {
        *llike_n = all_fail_llike;
        *ess_n = 0.0;
        for (int j = 0; j < J; j++) select[j] = (uint32_t)j;
        return 0;
    }
    const int P = next_pow2(T);
    double *a = buf, *b = buf + P;
    for (int t = 0; t < P; t++) {
        a[t] = 0.0;
        b[t] = 0.0;
        if (t < T && tmax[t] != -INFINITY) {
            double e = plom_exp(tmax[t] - m);
            a[t] = tsum[t] * e;
            b[t] = (tsum2[t] * e) * e;
        }
    }
    const double S = tree_sum(a, P), S2 = tree_sum(b, P);
    *llike_n = m + plom_log(S / (double)J);
    *ess_n = (S * S) / S2;
    /* fixed-point weights and their running integer prefix */
    uint64_t C = 0;
    for (int j = 0; j < J; j++) {
        uint64_t q = 0;
        if (logw[j] != -INFINITY) q = (uint64_t)((plom_exp(logw[j] - m) / S) * 0x1p62);
        C += q;
        cum[j] = C;
    }
    const uint64_t Q = C;
    const uint64_t unum = plom_resample_num(seed, stream, (uint32_t)n);
    const uint64_t J2 = (uint64_t)J << 33;
    /* the reference's two-pointer walk (smc.c:101-111), in exact integers */
    int i = 0;
    for (int k = 0; k < J; k++) {
        const u128 Rk = (u128)(((uint64_t)k << 33) + unum) * (u128)Q;
        while ((u128)cum[i] * J2 < Rk) i++;
        select[k] = (uint32_t)i;
    }
    return 1;
}

/* reference arithmetic (smc.c:27-112) on the same log-weights and uniform */
static int weight_select_reference(const double *logw, int J, int n, int n_ts, double like_min, uint32_t seed,
                                   uint32_t stream, double *llike_n, double *ess_n, uint32_t *select, double *w)
{
    double like_tot = 0.0, ess = 0.0;
    int nfail = 0;
    const double thr = pow(like_min, (double)n_ts);
    for (int j = 0; j < J; j++) {
        w[j] = exp(logw[j]);
        if (w[j] <= thr) {
            w[j] = 0.0;
            nfail++;
        } else {
            like_tot += w[j];
            ess += w[j] * w[j];
        }
    }
    if (nfail == J) {
        *llike_n = log(like_min) * n_ts;
        *ess_n = 0.0;
        for (int j = 0; j < J; j++) select[j] = (uint32_t)j;
        return 0;
    }
    for (int j = 0; j < J; j++) w[j] /= like_tot;
    *llike_n = log(like_tot / (double)J);
    *ess_n = (like_tot * like_tot) / ess;
    const double inc = 1.0 / (double)J;
    const double U = (double)plom_resample_num(seed, stream, (uint32_t)n) * 0x1p-33;
    double ran = inc * U;
    int i = 0;
    double cum = w[0];
    for (int k = 0; k < J; k++) {
        while (ran > cum && i < J - 1) {
            i++;
            cum += w[i];
        }
        select[k] = (uint32_t)i;
        ran += inc;
    }
    return 1;
}

/* ------------------------------------------------------------------ hat and residuals */
/* The per-interval summaries (compute_hat, hat.c:66-137; print_prediction_residuals,
 * print.c:632-701) in the build's arithmetic: fixed-point weights q_j (1 each when the
 * interval is not weighted), their total Q; the mean is the tree over tiles of the tile
 * trees of value x q, over Q; the 95 % bounds are found by sorting the values (ordered as
 * 64-bit keys) and walking the exact cumulative weights; the residuals use the offspring
 * counts of the selection, sums strided over 256 slots then a tree. */
static uint64_t hat_key(double x)
{
    uint64_t u;
    memcpy(&u, &x, 8);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

typedef struct { uint64_t key; uint64_t w; double v; } hat_item;

static int hat_cmp(const void *a, const void *b)
{
    const uint64_t x = ((const hat_item *)a)->key, y = ((const hat_item *)b)->key;
    return x < y ? -1 : (x > y);
}

static void hat_row(const oracle_model *M, const double *X, int J, const double *pv, const double *cvd,
                    const plom_drift_view *dm, double t, const uint64_t *q, uint64_t Q, double *row, double *vals,
                    hat_item *items, double *buf)
{
    const int NH = M->n_hat, S = M->S;
    const int T = (J + TILE - 1) / TILE, P = next_pow2(T);
    for (int j = 0; j < J; j++) M->hat_values(X + (size_t)j * S, pv, cvd, dm, t, vals + (size_t)j * NH);
    for (int k = 0; k < NH; k++) {
        double *tiles = buf + TILE;
        for (int tt = 0; tt < P; tt++) {
            tiles[tt] = 0.0;
            if (tt >= T) continue;
            for (int i = 0; i < TILE; i++) {
                const int j = tt * TILE + i;
                buf[i] = j < J ? vals[(size_t)j * NH + k] * (double)q[j] : 0.0;
            }
            tiles[tt] = tree_sum(buf, TILE);
        }
        row[3 * k + 1] = tree_sum(tiles, P) / (double)Q;
        for (int j = 0; j < J; j++) {
            items[j].v = vals[(size_t)j * NH + k];
            items[j].key = hat_key(items[j].v);
            items[j].w = q[j];
        }
        qsort(items, (size_t)J, sizeof(hat_item), hat_cmp);
        for (int b = 0; b < 2; b++) {
            const uint64_t p = b ? 39u : 1u;
            uint64_t cum = 0;
            int i = 0;
            for (; i < J; i++) {
                cum += items[i].w;
                if ((u128)cum * 40u >= (u128)Q * p) break;
            }
            row[3 * k + (b ? 2 : 0)] = items[i < J ? i : J - 1].v;
        }
    }
}

static void chan_tree(const double *c, const double *m, const double *q, int n, double *oc, double *om, double *oq);

/* The prediction residuals of interval n (print.c:632-701) over the resampled cloud (o_j
 * the copies of particle j), in the device's order: per tile of 256 particles W = sum o and,
 * per series, A = tree(o mu), B = tree(o var) (the tree of tree256), the tile mean A / W and
 * Q = tree(o (mu - A / W)^2); then pred = tree over the tiles of A, / J, var_obs = the same
 * of B, and var_state = M2 / (J - 1) with M2 the (count, mean, M2) merge of the tiles'
 * (W, A / W, Q) along the fixed pairwise tree (= sum o (mu - pred)^2). */
static void pred_res_row(const oracle_model *M, const double *X, int J, const double *pv, const double *cvd, double t,
                         const double *y, const uint32_t *select, uint32_t *offspring, double *row)
{
    const int S = M->S, NT = M->n_ts;
    const int T = (J + TILE - 1) / TILE, P = next_pow2(T);
    double *w = calloc((size_t)P * (size_t)(1 + 4 * NT), sizeof(double));   /* W, then per series A, B, mean, Q */
    double mu[TILE][16], var[TILE][16], obs[64], buf[TILE];
    uint32_t o[TILE];
    if (!w) { for (int k = 0; k < 2 * NT; k++) row[k] = NAN; return; }
    memset(offspring, 0, (size_t)J * sizeof(uint32_t));
    for (int k = 0; k < J; k++) offspring[select[k]]++;
    double *A = w + P, *B = A + (size_t)NT * P, *MT = B + (size_t)NT * P, *Q = MT + (size_t)NT * P;
    for (int tl = 0; tl < T; tl++) {
        for (int i = 0; i < TILE; i++) {
            const int j = tl * TILE + i;
            o[i] = j < J ? offspring[j] : 0u;
            for (int ts = 0; ts < NT; ts++) { mu[i][ts] = 0.0; var[i][ts] = 0.0; }
            if (o[i]) {
                M->proj2obs(X + (size_t)j * S, pv, cvd, obs);
                M->obs_moments(obs, pv, cvd, t, mu[i], var[i]);
            }
            buf[i] = (double)o[i];
        }
        const double W = tree_sum(buf, TILE);
        w[tl] = W;
        for (int ts = 0; ts < NT; ts++) {
            for (int i = 0; i < TILE; i++) buf[i] = o[i] ? (double)o[i] * mu[i][ts] : 0.0;
            const double a = tree_sum(buf, TILE);
            for (int i = 0; i < TILE; i++) buf[i] = o[i] ? (double)o[i] * var[i][ts] : 0.0;
            const double bb = tree_sum(buf, TILE);
            const double mt = W > 0.0 ? a / W : 0.0;
            for (int i = 0; i < TILE; i++) {
                const double d = mu[i][ts] - mt;
                buf[i] = o[i] ? (double)o[i] * (d * d) : 0.0;
            }
            A[(size_t)ts * P + tl] = a;
            B[(size_t)ts * P + tl] = bb;
            MT[(size_t)ts * P + tl] = mt;
            Q[(size_t)ts * P + tl] = tree_sum(buf, TILE);
        }
    }
    for (int ts = 0; ts < NT; ts++) {
        const double pred = tree_sum(A + (size_t)ts * P, P) / (double)J;
        const double vobs = tree_sum(B + (size_t)ts * P, P) / (double)J;
        double c, m, q;
        chan_tree(w, MT + (size_t)ts * P, Q + (size_t)ts * P, P, &c, &m, &q);
        const double vstate = q / (double)(J - 1);
        row[2 * ts] = pred;
        row[2 * ts + 1] = (y[ts] - pred) / plom_sqrt(vstate + vobs);
    }
    free(w);
}

/* ------------------------------------------------------------------ the filter */
int oracle_smc(int model, int J, int N, uint32_t noises_off, uint32_t flags, uint32_t seed, uint32_t stream,
               double dt, double like_min, const double *times, const double *data, const double *pv,
               const double *x0, int mode, int nthreads,
               double *loglik, int32_t *n_all_fail, double *loglik_n, double *ess_n, int32_t *ancestry,
               double *final_state, double *logw_out, double *elapsed_s,
               int32_t *lin_slots, double *lin_state, double *all_states, double *hat, double *pred_res,
               double *lin_obs, const int32_t *cov_n, const double *cov_x, const double *cov_y, double eps_abs,
               double eps_rel, const int32_t *dm_group, const int32_t *dm_tr, const int32_t *dm_sc,
               const int32_t *dm_dur, const double *dm_min, const double *dm_max, const double *dm_mult,
               const int32_t *dm_vol)
{
    if (model < 0 || model >= N_MODELS || J < 1 || N < 0) return -1;
    const oracle_model *M = &MODELS[model];
    drift_arg da;
    drift_build(M, dm_group, dm_tr, dm_sc, dm_dur, dm_min, dm_max, dm_mult, dm_vol, &da);
    const plom_drift_view *dm = drift_view(&da);
    cov_view cv;
    const int crc = cov_build(M, N, times, dt, cov_n, cov_x, cov_y, pv, &cv);
    if (crc) { cov_free(&cv); return crc; }
    const int S = M->S, n_ts = M->n_ts;
    const int T = (J + TILE - 1) / TILE;
    double *X = malloc((size_t)J * S * sizeof(double));
    double *Xn = malloc((size_t)J * S * sizeof(double));
    double *logw = malloc((size_t)J * sizeof(double));
    double *work = malloc((size_t)(3 * T + 4 * next_pow2(T) + 2 * TILE + J) * sizeof(double));
    uint64_t *cum = malloc((size_t)J * sizeof(uint64_t));
    uint32_t *select = malloc((size_t)J * sizeof(uint32_t));
    const int lineage = lin_slots != NULL && lin_state != NULL && N > 0;
    int32_t *anc = ancestry;
    if (lineage && !anc) anc = malloc((size_t)N * J * sizeof(int32_t));
    if (!X || !Xn || !logw || !work || !cum || !select || (lineage && !anc)) return -2;
    for (int j = 0; j < J; j++) memcpy(X + (size_t)j * S, x0, (size_t)S * sizeof(double));
    if (all_states) memcpy(all_states, X, (size_t)J * S * sizeof(double));
    int last_resampled = 0;
    const int NH = M->n_hat;
    double *hvals = NULL, *hbuf = NULL;
    hat_item *hitems = NULL;
    uint64_t *hq = NULL;
    uint32_t *offspring = NULL;
    if (hat) {
        hvals = malloc((size_t)J * NH * sizeof(double));
        hbuf = malloc((size_t)(TILE + next_pow2(T)) * sizeof(double));
        hitems = malloc((size_t)J * sizeof(hat_item));
        hq = malloc((size_t)J * sizeof(uint64_t));
        offspring = malloc((size_t)J * sizeof(uint32_t));
        if (!hvals || !hbuf || !hitems || !hq || !offspring || M->n_ts > 16) return -2;
        double v0[256];
        M->hat_values(x0, pv, cv_at(&cv, 0), dm, times[0], v0);
        for (int k = 0; k < NH; k++) hat[3 * k] = hat[3 * k + 1] = hat[3 * k + 2] = v0[k];
    }
    const int filter = !(flags & FLAG_NO_FILTER);
    const double all_fail_llike = plom_log(like_min) * (double)n_ts;
    double llike = 0.0;
    int naf = 0;
    struct timespec ts0, ts1;
    clock_gettime(CLOCK_MONOTONIC, &ts0);
    for (int n = 0; n < N; n++) {
        int nonan = 0;
        for (int k = 0; k < n_ts; k++) nonan += !isnan(data[(size_t)n * n_ts + k]);
        prop_job job;
        memset(&job, 0, sizeof job);
        job.M = M; job.X = X; job.logw = logw; job.pv = pv; job.cv = &cv; job.dm = dm; job.times = times; job.data = data;
        job.J = J; job.n = n; job.nonan = nonan; job.filter = filter; job.n_ts = n_ts;
        job.seed = seed; job.stream = stream; job.noises_off = noises_off;
        job.dt = dt; job.sqdt = plom_sqrt(dt); job.like_min = like_min; job.mode = mode;
        job.eps_abs = eps_abs; job.eps_rel = eps_rel;
        propagate_all(&job, nthreads);
        if (all_states) memcpy(all_states + (size_t)(n + 1) * J * S, X, (size_t)J * S * sizeof(double));
        double ln = 0.0, es = 0.0;
        int resampled = 0;
        if (filter && nonan) {
            if (mode == ORACLE_CANONICAL)
                resampled = weight_select_canonical(logw, J, n, seed, stream, all_fail_llike, &ln, &es, select, work, cum);
            else
                resampled = weight_select_reference(logw, J, n, n_ts, like_min, seed, stream, &ln, &es, select, work);
            if (!resampled) naf++;
            llike += ln;
            if (logw_out) memcpy(logw_out + (size_t)n * J, logw, (size_t)J * sizeof(double));
        } else {
            for (int j = 0; j < J; j++) select[j] = (uint32_t)j;
        }
        if (loglik_n) loglik_n[n] = ln;
        if (ess_n) ess_n[n] = es;
        last_resampled = resampled;
        if (hat) {
            uint64_t Q = (uint64_t)J;
            if (resampled && mode == ORACLE_CANONICAL) {
                for (int j = 0; j < J; j++) hq[j] = cum[j] - (j ? cum[j - 1] : 0ull);
                Q = cum[J - 1];
            } else {
                for (int j = 0; j < J; j++) hq[j] = 1;
            }
            hat_row(M, X, J, pv, cv_at(&cv, n + 1), dm, times[n + 1], hq, Q, hat + (size_t)(n + 1) * NH * 3, hvals,
                    hitems, hbuf);
            double *pr = pred_res + (size_t)n * n_ts * 2;
            if (filter && nonan)
                pred_res_row(M, X, J, pv, cv_at(&cv, n + 1), times[n + 1], data + (size_t)n * n_ts, select, offspring, pr);
            else
                for (int k = 0; k < 2 * n_ts; k++) pr[k] = NAN;
        }
        if (anc)
            for (int j = 0; j < J; j++) anc[(size_t)n * J + j] = (int32_t)select[j];
        /* resample_X: gather the ancestors' states */
        for (int j = 0; j < J; j++) memcpy(Xn + (size_t)j * S, X + (size_t)select[j] * S, (size_t)S * sizeof(double));
        double *tmp = X; X = Xn; Xn = tmp;
    }
    clock_gettime(CLOCK_MONOTONIC, &ts1);
    if (elapsed_s) *elapsed_s = (double)(ts1.tv_sec - ts0.tv_sec) + 1e-9 * (double)(ts1.tv_nsec - ts0.tv_nsec);
    if (final_state)
        for (int k = 0; k < S; k++)
            for (int j = 0; j < J; j++) final_state[(size_t)k * J + j] = X[(size_t)j * S + k];
    *loglik = llike;
    *n_all_fail = naf;
    if (lineage) {
        /* the final particle: inverse CDF of the last interval's weights at one uniform
         * (print.c:730-737), in the arithmetic of the mode; equal weights without them */
        const uint64_t unum = plom_lineage_num(seed, stream);
        int j = 0;
        if (last_resampled && mode == ORACLE_CANONICAL) {
            const uint64_t Q = cum[J - 1];
            while ((u128)cum[j] << 33 < (u128)unum * Q) j++;
        } else if (last_resampled) {
            const double U = (double)unum * 0x1p-33;
            double cw = work[0];
            while (cw < U && j < J - 1) cw += work[++j];
        } else {
            j = (int)((unum * (uint64_t)J + (1ull << 33) - 1ull) >> 33) - 1;
        }
        if (j < 0) j = 0;
        if (j >= J) j = J - 1;
        /* its ancestors, then its path regenerated with the same noise */
        lin_slots[N - 1] = j;
        for (int n = N - 1; n >= 1; n--) lin_slots[n - 1] = anc[(size_t)(n - 1) * J + lin_slots[n]];
        double x[256];
        memcpy(x, x0, (size_t)S * sizeof(double));
        memcpy(lin_state, x, (size_t)S * sizeof(double));
        for (int n = 0; n < N; n++) {
            plom_nstream64 ns;
            plom_nstream64_init(&ns, seed, stream, (uint32_t)n, (uint32_t)lin_slots[n], plom_icdf_tab);
            plom_psr_tables(&ns, plom_bm_tab, plom_lfact_tab);
            M->interval(x, pv, &ns, times[n], times[n + 1], dt, plom_sqrt(dt), noises_off, cv_steps(&cv, n), cv.cs, eps_abs,
                        eps_rel, dm, cv_splines(&cv, noises_off));
            memcpy(lin_state + (size_t)(n + 1) * S, x, (size_t)S * sizeof(double));
        }
        if (lin_obs)
            for (int n = 0; n <= N; n++) {
                double obs[64], mu[64], var[64];
                M->proj2obs(lin_state + (size_t)n * S, pv, cv_at(&cv, n), obs);
                M->obs_moments(obs, pv, cv_at(&cv, n), times[n], mu, var);
                plom_lineage_obs(seed, stream, (uint32_t)n, n_ts, mu, var, plom_bm_tab, lin_obs + (size_t)n * n_ts * 2);
            }
    }
    if (anc && anc != ancestry) free(anc);
    free(hvals); free(hbuf); free(hitems); free(hq); free(offspring);
    free(X); free(Xn); free(logw); free(work); free(cum); free(select);
    cov_free(&cv);
    return 0;
}

/* ------------------------------------------------------------------ iterated filtering */
/* recursive contiguous pairwise merge of n (count, mean, M2) triples (n a power of two) */
static void chan_tree(const double *c, const double *m, const double *q, int n, double *oc, double *om, double *oq)
{
    if (n == 1) { *oc = c[0]; *om = m[0]; *oq = q[0]; return; }
    double ca, ma, qa, cb, mb, qb;
    chan_tree(c, m, q, n / 2, &ca, &ma, &qa);
    chan_tree(c + n / 2, m + n / 2, q + n / 2, n / 2, &cb, &mb, &qb);
    plom_mif_combine(ca, ma, qa, cb, mb, qb, oc, om, oq);
}

/* One pass of iterated filtering with per-particle parameters.  theta0 [n_theta][J] and
 * x0 [S][J] are the particles' starting parameters and states; after an interval n-1 with
 * data, every component with sd > 0 moves by (mut_fac[n-1] * sd) * z (z: plom_mif_normals of
 * the particle's new slot); the pv slots s < n_pp with pv_src[s] >= 0 are the natural values
 * of their components.  Outputs as the filter's, plus mif_mean / mif_var rows (n+1) for the
 * intervals with data and, optionally, every particle's theta per row [(N+1)][n_theta][J]. */
int oracle_mif_pass(int model, int J, int N, uint32_t noises_off, uint32_t seed, uint32_t stream, double dt,
                    double like_min, const double *times, const double *data, const double *pv, int n_theta,
                    const int32_t *pv_src, const int32_t *tr, const int32_t *sc, const int32_t *dur, const double *mn,
                    const double *mx, const double *mult, const double *sd, const double *mut_fac,
                    const double *theta0, const double *x0, int nthreads, double *loglik, int32_t *n_all_fail,
                    double *loglik_n, double *ess_n, double *mif_mean, double *mif_var, double *theta_out,
                    double *logw_out, int32_t *ancestry, const int32_t *cov_n, const double *cov_x,
                    const double *cov_y, double eps_abs, double eps_rel, const double *chol, const int32_t *prior,
                    const double *prior_pow, const int32_t *prior_until, const int32_t *dm_group, const int32_t *dm_tr,
                    const int32_t *dm_sc, const int32_t *dm_dur, const double *dm_min, const double *dm_max,
                    const double *dm_mult, const int32_t *dm_vol)
{
    if (model < 0 || model >= N_MODELS || J < 1 || N < 0 || n_theta < 1) return -1;
    if (chol && n_theta > PLOM_MIF_MVN_MAX) return -1;
    const oracle_model *M = &MODELS[model];
    /* drifting parameters: every particle's volatility is its own pv row's slot */
    drift_arg da;
    drift_build(M, dm_group, dm_tr, dm_sc, dm_dur, dm_min, dm_max, dm_mult, dm_vol, &da);
    cov_view cv;
    const int crc = cov_build(M, N, times, dt, cov_n, cov_x, cov_y, pv, &cv);
    if (crc) { cov_free(&cv); return crc; }
    const int S = M->S, n_ts = M->n_ts, P = M->n_pp;
    const int T = (J + TILE - 1) / TILE, PT = next_pow2(T);
    double *X = malloc((size_t)J * S * sizeof(double)), *Xn = malloc((size_t)J * S * sizeof(double));
    double *TH = malloc((size_t)J * n_theta * sizeof(double)), *THn = malloc((size_t)J * n_theta * sizeof(double));
    double *logw = malloc((size_t)J * sizeof(double));
    double *work = malloc((size_t)(3 * T + 4 * PT + 2 * TILE + J) * sizeof(double));
    uint64_t *cum = malloc((size_t)J * sizeof(uint64_t));
    uint32_t *select = malloc((size_t)J * sizeof(uint32_t));
    double *buf = malloc((size_t)TILE * sizeof(double));
    double *tc = malloc((size_t)PT * 5 * sizeof(double));
    if (!X || !Xn || !TH || !THn || !logw || !work || !cum || !select || !buf || !tc) return -2;
    for (int j = 0; j < J; j++) {
        for (int k = 0; k < S; k++) X[(size_t)j * S + k] = x0[(size_t)k * J + j];
        for (int o = 0; o < n_theta; o++) TH[(size_t)j * n_theta + o] = theta0[(size_t)o * J + j];
    }
    if (theta_out) memcpy(theta_out, theta0, (size_t)n_theta * J * sizeof(double));
    /* per-particle parameter rows: the caller's pv, the per-particle slots overwritten */
    const int pvn = M->pv_size;
    double *PVJ = malloc((size_t)J * pvn * sizeof(double));
    double *LP = prior ? malloc((size_t)J * sizeof(double)) : NULL;
    if (!PVJ || (prior && !LP)) return -2;
    const double all_fail_llike = plom_log(like_min) * (double)n_ts;
    double llike = 0.0;
    int naf = 0;
    int prev_data = 0;
    for (int n = 0; n < N; n++) {
        int nonan = 0;
        for (int k = 0; k < n_ts; k++) nonan += !isnan(data[(size_t)n * n_ts + k]);
        /* the random walk after an interval with data, then the natural parameters */
        const double fac = (n > 0 && prev_data) ? mut_fac[n - 1] : 0.0;
        for (int j = 0; j < J; j++) {
            double *th = TH + (size_t)j * n_theta;
            if (chol) {
                /* correlated walk: th_o += fac sum_{q<=o} L[o][q] z_q */
                double z[PLOM_MIF_MVN_MAX];
                for (int o = 0; o < n_theta; o += 2) {
                    double z0 = 0.0, z1 = 0.0;
                    if (fac != 0.0)
                        plom_mif_normals(seed, stream, (uint32_t)n, (uint32_t)j, (uint32_t)(o >> 1), plom_bm_tab, &z0, &z1);
                    z[o] = z0;
                    if (o + 1 < n_theta) z[o + 1] = z1;
                }
                for (int o = 0; o < n_theta && fac != 0.0; o++) {
                    double acc = 0.0;
                    for (int q = 0; q <= o; q++) acc = acc + chol[(size_t)o * n_theta + q] * z[q];
                    th[o] = th[o] + fac * acc;
                }
            } else
            for (int o = 0; o < n_theta; o += 2) {
                const double sd0 = fac * sd[o];
                const double sd1 = (o + 1 < n_theta) ? fac * sd[o + 1] : 0.0;
                double z0 = 0.0, z1 = 0.0;
                if (sd0 != 0.0 || sd1 != 0.0)
                    plom_mif_normals(seed, stream, (uint32_t)n, (uint32_t)j, (uint32_t)(o >> 1), plom_bm_tab, &z0, &z1);
                if (sd0 != 0.0) th[o] = th[o] + sd0 * z0;
                if (o + 1 < n_theta && sd1 != 0.0) th[o + 1] = th[o + 1] + sd1 * z1;
            }
            double *row = PVJ + (size_t)j * pvn;
            memcpy(row, pv, (size_t)pvn * sizeof(double));
            for (int s2 = 0; s2 < P; s2++) {
                const int o = pv_src[s2];
                if (o >= 0) row[s2] = plom_mif_natural(tr[o], sc[o], dur[o], th[o], mn[o], mx[o], mult[o]);
            }
            /* mif --prior: the particle's prior terms at this interval, in component order */
            if (prior) {
                double pt = 0.0;
                for (int o = 0; o < n_theta; o++)
                    if (prior[o] >= 0 && n < prior_until[o])
                        pt = pt + plom_mif_prior_term(prior[o], prior_pow[o], tr[o], th[o], mn[o], mx[o], like_min);
                LP[j] = pt;
            }
        }
        if (theta_out)
            for (int o = 0; o < n_theta; o++)
                for (int j = 0; j < J; j++)
                    theta_out[((size_t)(n + 1) * n_theta + o) * J + j] = TH[(size_t)j * n_theta + o];
        prop_job job;
        memset(&job, 0, sizeof job);
        job.M = M; job.X = X; job.logw = logw; job.pv = PVJ; job.pv_stride = (size_t)pvn; job.cv = &cv; job.times = times;
        job.data = data; job.J = J; job.n = n; job.nonan = nonan; job.filter = 1; job.n_ts = n_ts;
        job.seed = seed; job.stream = stream; job.noises_off = noises_off;
        job.dt = dt; job.sqdt = plom_sqrt(dt); job.like_min = like_min; job.mode = ORACLE_CANONICAL;
        job.eps_abs = eps_abs; job.eps_rel = eps_rel;
        job.lprior = LP; job.all_fail_llike = all_fail_llike;
        job.dm = drift_view(&da);
        propagate_all(&job, nthreads);
        double ln = 0.0, es = 0.0;
        int resampled = 0;
        if (nonan) {
            resampled = weight_select_canonical(logw, J, n, seed, stream, all_fail_llike, &ln, &es, select, work, cum);
            if (!resampled) naf++;
            llike += ln;
            if (logw_out) memcpy(logw_out + (size_t)n * J, logw, (size_t)J * sizeof(double));
            /* moments of theta: weights exp(l - m) (1 each when every particle failed) */
            double m = -INFINITY;
            for (int j = 0; j < J; j++) if (logw[j] > m) m = logw[j];
            double *tw = tc, *ts1 = tc + PT, *tcn = tc + 2 * PT, *tmn = tc + 3 * PT, *tq = tc + 4 * PT;
            for (int t = 0; t < PT; t++) tw[t] = 0.0;
            for (int t = 0; t < T; t++) {
                for (int i = 0; i < TILE; i++) {
                    const int j = t * TILE + i;
                    double w = 0.0;
                    if (j < J) w = resampled ? (logw[j] == -INFINITY ? 0.0 : plom_exp(logw[j] - m)) : 1.0;
                    buf[i] = w;
                }
                tw[t] = tree_sum(buf, TILE);
            }
            const double W = tree_sum(tw, PT);
            for (int o = 0; o < n_theta; o++) {
                for (int t = 0; t < PT; t++) { ts1[t] = 0.0; tcn[t] = 0.0; tmn[t] = 0.0; tq[t] = 0.0; }
                for (int t = 0; t < T; t++) {
                    const int rem = J - t * TILE;
                    const double ct = (double)(rem < TILE ? rem : TILE);
                    for (int i = 0; i < TILE; i++) {
                        const int j = t * TILE + i;
                        double w = 0.0;
                        if (j < J) w = resampled ? (logw[j] == -INFINITY ? 0.0 : plom_exp(logw[j] - m)) : 1.0;
                        buf[i] = j < J ? w * TH[(size_t)j * n_theta + o] : 0.0;
                    }
                    ts1[t] = tree_sum(buf, TILE);
                    for (int i = 0; i < TILE; i++) {
                        const int j = t * TILE + i;
                        buf[i] = j < J ? TH[(size_t)j * n_theta + o] : 0.0;
                    }
                    const double mt = tree_sum(buf, TILE) / ct;
                    for (int i = 0; i < TILE; i++) {
                        const int j = t * TILE + i;
                        const double d = j < J ? TH[(size_t)j * n_theta + o] - mt : 0.0;
                        buf[i] = d * d;
                    }
                    tcn[t] = ct;
                    tmn[t] = mt;
                    tq[t] = tree_sum(buf, TILE);
                }
                double cc, cm, cq;
                chan_tree(tcn, tmn, tq, PT, &cc, &cm, &cq);
                mif_mean[(size_t)(n + 1) * n_theta + o] = tree_sum(ts1, PT) / W;
                mif_var[(size_t)(n + 1) * n_theta + o] = cq / (double)(J - 1);
            }
        } else {
            for (int j = 0; j < J; j++) select[j] = (uint32_t)j;
            for (int o = 0; o < n_theta; o++) {
                mif_mean[(size_t)(n + 1) * n_theta + o] = NAN;
                mif_var[(size_t)(n + 1) * n_theta + o] = NAN;
            }
        }
        if (loglik_n) loglik_n[n] = ln;
        if (ess_n) ess_n[n] = es;
        prev_data = nonan > 0;
        if (ancestry)
            for (int j = 0; j < J; j++) ancestry[(size_t)n * J + j] = (int32_t)select[j];
        /* resample states and parameters together */
        for (int j = 0; j < J; j++) {
            memcpy(Xn + (size_t)j * S, X + (size_t)select[j] * S, (size_t)S * sizeof(double));
            memcpy(THn + (size_t)j * n_theta, TH + (size_t)select[j] * n_theta, (size_t)n_theta * sizeof(double));
        }
        double *tmp = X; X = Xn; Xn = tmp;
        tmp = TH; TH = THn; THn = tmp;
    }
    for (int o = 0; o < n_theta; o++) { mif_mean[o] = NAN; mif_var[o] = NAN; }
    *loglik = llike;
    *n_all_fail = naf;
    free(X); free(Xn); free(TH); free(THn); free(logw); free(work); free(cum); free(select); free(buf); free(tc);
    free(PVJ);
    free(LP);
    cov_free(&cv);
    return 0;
}

/* one particle through N intervals, no filtering: the state at every observation time,
 * out[(n+1)*S + k]; out[0..S) = x0 (with incidences reset) */
int oracle_trajectory(int model, int N, uint32_t noises_off, uint32_t seed, uint32_t stream, uint32_t j,
                      double dt, const double *times, const double *pv, const double *x0, double *out,
                      const int32_t *cov_n, const double *cov_x, const double *cov_y, double eps_abs, double eps_rel,
                      const int32_t *dm_group, const int32_t *dm_tr, const int32_t *dm_sc, const int32_t *dm_dur,
                      const double *dm_min, const double *dm_max, const double *dm_mult, const int32_t *dm_vol)
{
    if (model < 0 || model >= N_MODELS) return -1;
    const oracle_model *M = &MODELS[model];
    drift_arg da;
    drift_build(M, dm_group, dm_tr, dm_sc, dm_dur, dm_min, dm_max, dm_mult, dm_vol, &da);
    cov_view cv;
    const int crc = cov_build(M, N, times, dt, cov_n, cov_x, cov_y, pv, &cv);
    if (crc) { cov_free(&cv); return crc; }
    const int S = M->S;
    double x[256];
    memcpy(x, x0, (size_t)S * sizeof(double));
    memcpy(out, x0, (size_t)S * sizeof(double));
    for (int n = 0; n < N; n++) {
        plom_nstream64 ns;
        plom_nstream64_init(&ns, seed, stream, (uint32_t)n, j, plom_icdf_tab);
        plom_psr_tables(&ns, plom_bm_tab, plom_lfact_tab);
        M->interval(x, pv, &ns, times[n], times[n + 1], dt, plom_sqrt(dt), noises_off, cv_steps(&cv, n), cv.cs, eps_abs,
                    eps_rel, drift_view(&da), cv_splines(&cv, noises_off));
        memcpy(out + (size_t)(n + 1) * S, x, (size_t)S * sizeof(double));
    }
    cov_free(&cv);
    return 0;
}

/* observed quantities and log-likelihood of one state (for tests) */
int oracle_obs_loglik(int model, const double *x, const double *y, const double *pv, double t, double like_min,
                      double *obs, double *loglik, int32_t *n_floor)
{
    if (model < 0 || model >= N_MODELS) return -1;
    const oracle_model *M = &MODELS[model];
    int nf = 0;
    M->proj2obs(x, pv, pv + M->fix_off, obs);
    *loglik = M->loglik(obs, y, pv, pv + M->fix_off, t, like_min, &nf);
    *n_floor = nf;
    return 0;
}

/* ------------------------------------------------------------------ the RKF45 integrator */
/* analytic right-hand sides for the tests of csrc/plom_ode_impl.h:
 *   kind 0   y' = p0 y (dimension 1)
 *   kind 1   y' = p0 + p1 t + p2 t^2 + p3 t^3 + p4 t^4 (dimension 1)
 *   kind 2   y0' = y1, y1' = -p0 y0 (harmonic oscillator, dimension 2) */
static int ode_kind;
static void ode_test_rhs(double t, const double *y, const double *pv, const double *lp, const double *cv, double *f)
{
    (void)lp; (void)cv;
    if (ode_kind == 0) {
        f[0] = pv[0] * y[0];
        f[1] = 0.0;
    } else if (ode_kind == 1) {
        f[0] = pv[0] + t * (pv[1] + t * (pv[2] + t * (pv[3] + t * pv[4])));
        f[1] = 0.0;
    } else {
        f[0] = y[1];
        f[1] = -pv[0] * y[0];
    }
}

static int ode_test_count;
static void ode_test_rhs_counted(double t, const double *y, const double *pv, const double *lp, const double *cv,
                                 const plom_spline_set *sp, double *f)
{
    (void)sp;
    ode_test_count++;
    ode_test_rhs(t, y, pv, lp, cv, f);
}

#define PLOM_ODE_FN ode_test_integrate
#define PLOM_ODE_RHS ode_test_rhs_counted
#define PLOM_ODE_N 2
#define PLOM_ODE_QUAL static
#include "plom_ode_impl.h"

/* integrate a test system (kind above, parameters p[5]) over [t0, t1] from y (2 values; the
 * second is carried along for the 1-d kinds), first step h; returns the right-hand side
 * evaluations in *n_eval, and 0 or -1 (step bound) */
int oracle_ode_test(int kind, const double *p, double *y, double t0, double t1, double h, double eps_abs, double eps_rel,
                    int32_t *n_eval)
{
    ode_kind = kind;
    ode_test_count = 0;
    const int rc = ode_test_integrate(y, p, NULL, NULL, NULL, t0, t1, h, eps_abs, eps_rel);
    *n_eval = ode_test_count;
    return rc;
}

/* the observation mean of state x at time t and one observation drawn for (row n, slot)
 * (plom_obs_draw); cvd: the covariates at t (NULL: those of pv) */
int oracle_obs_draw(int model, const double *x, const double *pv, const double *cvd, double t, uint32_t seed,
                    uint32_t stream, uint32_t n, uint32_t slot, double *out)
{
    if (model < 0 || model >= N_MODELS) return -1;
    const oracle_model *M = &MODELS[model];
    double obs[64], mu[64], var[64];
    const double *cv = cvd ? cvd : pv + M->fix_off;
    M->proj2obs(x, pv, cv, obs);
    M->obs_moments(obs, pv, cv, t, mu, var);
    plom_obs_draw(seed, stream, n, slot, M->n_ts, mu, var, plom_bm_tab, out);
    return 0;
}

/* shared scalar arithmetic, host build (compared with the device build bit for bit) */
void oracle_math(int n, const double *in, double *out)
{
    for (int i = 0; i < n; i++) {
        const double x = in[i];
        out[i * 8 + 0] = plom_exp(x);
        out[i * 8 + 1] = plom_log(x);
        out[i * 8 + 2] = plom_Phi(x);
        out[i * 8 + 3] = plom_sin(x);
        out[i * 8 + 4] = plom_cos(x);
        out[i * 8 + 5] = plom_sqrt(x);
        out[i * 8 + 6] = plom_sqrt_mid(x);
        out[i * 8 + 7] = plom_sqrt_rate(x);
    }
}

/* the first n Gaussians (Box-Muller pairs in order) of particle j over interval n0 */
void oracle_normals(uint32_t seed, uint32_t stream, uint32_t n0, uint32_t j, int n, double *out)
{
    plom_nstream64 ns;
    plom_nstream64_init(&ns, seed, stream, n0, j, plom_icdf_tab);
    plom_psr_tables(&ns, plom_bm_tab, plom_lfact_tab);
    for (int i = 0; i + 1 < n; i += 2) plom_normal_pair64(&ns, &out[i], &out[i + 1]);
    if (n & 1) {
        double spare;
        plom_normal_pair64(&ns, &out[n - 1], &spare);
    }
}

/* the Box-Muller transform of given word pairs (xy[2i], xy[2i+1]) -> out[2i], out[2i+1] */
void oracle_bm_pairs(int n, const uint32_t *xy, double *out)
{
    for (int i = 0; i < n; i++) plom_bm_pair(xy[2 * i], xy[2 * i + 1], plom_bm_tab, &out[2 * i], &out[2 * i + 1]);
}

/* quotients x[i] / c[i] from the divisor's reciprocal (plom_div_rc, the generated steps'
 * division by a covariate) */
void oracle_div_rc(int n, const double *x, const double *c, double *out)
{
    for (int i = 0; i < n; i++) out[i] = plom_div_rc(x[i], c[i], plom_recip_uniform(c[i]));
}

/* the stream's Gaussian transform of the 32-bit words w[0..n) (plom_icdf_normal) */
void oracle_icdf(int n, const uint32_t *w, double *out)
{
    for (int i = 0; i < n; i++) out[i] = plom_icdf_normal(w[i], plom_icdf_tab);
}

/* terms_forcing(amp) at the times t[0..n) (days) for one population's school record rec
 * (plom_terms_forcing) */
void oracle_terms_forcing(int n, double amp, const double *t, const double *rec, double *out)
{
    for (int i = 0; i < n; i++) out[i] = plom_terms_forcing(amp, t[i], rec);
}

/* the psr samplers' table logarithm of x[0..n) (plom_log_tab) */
void oracle_log_tab(int n, const double *x, double *out)
{
    for (int i = 0; i < n; i++) out[i] = plom_log_tab(x[i], plom_bm_tab);
}

void oracle_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out) { plom_philox4x32_10(ctr, key, out); }

uint32_t oracle_xoshiro(uint32_t *state, int n, uint32_t *out)
{
    plom_xoshiro128 r = {state[0], state[1], state[2], state[3]};
    for (int i = 0; i < n; i++) out[i] = plom_xoshiro_next(&r);
    state[0] = r.s0; state[1] = r.s1; state[2] = r.s2; state[3] = r.s3;
    return 0;
}

/* weighting + selection of one interval on given log-weights (for tests); returns 1 if
 * resampled, 0 for the all-fail branch */
int oracle_resample(int J, const double *logw, int n, int n_ts, double like_min, uint32_t seed, uint32_t stream,
                    int mode, int32_t *select_out, double *llike_n, double *ess_n)
{
    const int T = (J + TILE - 1) / TILE;
    double *work = malloc((size_t)(3 * T + 4 * next_pow2(T) + 2 * TILE + J) * sizeof(double));
    uint64_t *cum = malloc((size_t)J * sizeof(uint64_t));
    uint32_t *sel = malloc((size_t)J * sizeof(uint32_t));
    int r;
    if (mode == ORACLE_CANONICAL)
        r = weight_select_canonical(logw, J, n, seed, stream, plom_log(like_min) * (double)n_ts, llike_n, ess_n, sel, work, cum);
    else
        r = weight_select_reference(logw, J, n, n_ts, like_min, seed, stream, llike_n, ess_n, sel, work);
    for (int j = 0; j < J; j++) select_out[j] = (int32_t)sel[j];
    free(work); free(cum); free(sel);
    return r;
}

uint64_t oracle_resample_num(uint32_t seed, uint32_t stream, uint32_t n) { return plom_resample_num(seed, stream, n); }
