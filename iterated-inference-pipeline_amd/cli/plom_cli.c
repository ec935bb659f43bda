/*
 * plom_cli.c -- host-side logic shared by the `smc` and `pmcmc` commands (plom_cli.h).
 *
 * Parameter handling follows the reference: routers from partitions
 * (build.c:267-475), transform then back-transform with scale, unit multiplier and
 * duration->rate (transform.c:23-107, 185-301, 332-410, 462-514), the theta remainder
 * (json.c:405-424, proposal.c:87-114), initial state = proportions x population size at
 * t = 0 (prediction_util.c:82-115, build.c:1179-1185).
 */
#include "plom_cli.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ZERO_LOG 1e-17
#define ONE_LOGIT 0.999999999
#define BUFSIZE (5000 * 1024)

int CLI_QUIET = 0;
int CLI_TRACE_17G = 0;

void cli_log(const char *msg)
{
    if (!CLI_QUIET) printf("%s\n", msg);
}

void cli_die(const char *msg)
{
    fprintf(stderr, "error: %s\n", msg);
    exit(EXIT_FAILURE);
}

int cli_implementation(int argc, char **argv, const char *help)
{
    if (argc == 0 || !strcmp(argv[0], "psr")) return CLI_PSR;
    if (!strcmp(argv[0], "sde")) return CLI_SDE;
    if (!strcmp(argv[0], "ode")) return CLI_ODE;
    fprintf(stderr, "%s", help);
    exit(1);
}

void cli_apply_implementation(int impl, int fp32, unsigned int *noises_off, uint32_t *flags)
{
    if (impl != CLI_SDE && fp32) cli_die("the psr and ode implementations run in FP64 only (use 'sde' with --fp32)");
    if (impl == CLI_PSR) *flags |= PLOM_FLAG_PSR;
    if (impl == CLI_ODE) {
        *flags |= PLOM_FLAG_ODE;
        /* 'ode' turns every noise off (C/smc/main.c:227-229) */
        *noises_off |= PLOM_NO_DEM_STO_BIT | PLOM_NO_ENV_STO_BIT | PLOM_NO_DRIFT_BIT;
    }
}

/* ------------------------------------------------------------------ transforms */
double cli_f_tr(cli_tr k, double x, double a, double b)
{
    double safe, ratio;
    switch (k) {
    case TR_LOG:
        safe = (x > ZERO_LOG) ? x : ZERO_LOG;
        return log(safe);
    case TR_LOGIT:
        safe = (x > ZERO_LOG) ? x : ZERO_LOG;
        safe = (safe < ONE_LOGIT) ? safe : ONE_LOGIT;
        return log(safe / (1.0 - safe));
    case TR_LOGIT_AB:
        if (a == b) return x;
        ratio = (x - a) / (b - x);
        if (ratio < ZERO_LOG) ratio = ZERO_LOG;
        else if (ratio > (1.0 / ZERO_LOG)) ratio = 1.0 / ZERO_LOG;
        return log(ratio);
    default:
        return x;
    }
}

double cli_f_inv(cli_tr k, double x, double a, double b)
{
    switch (k) {
    case TR_LOG:
        return exp(x);
    case TR_LOGIT:
        return (x > 0) ? 1.0 / (1.0 + exp(-x)) : exp(x) / (1.0 + exp(x));
    case TR_LOGIT_AB:
        if (a == b) return x;
        return (x < 0) ? (b * exp(x) + a) / (1.0 + exp(x)) : (b + a * exp(-x)) / (1.0 + exp(-x));
    default:
        return x;
    }
}

/* derivative of f_inv (transform.c:121-175); the identity's entry is f_id itself, as there */
double cli_f_inv_der(cli_tr k, double x, double a, double b)
{
    double e;
    switch (k) {
    case TR_LOG:
        return exp(x);
    case TR_LOGIT:
        e = exp(x > 0 ? -x : x);
        return e / ((1.0 + e) * (1.0 + e));
    case TR_LOGIT_AB:
        if (a == b) return x;
        e = exp(x > 0 ? -x : x);
        return (b - a) * e / ((e + 1.0) * (e + 1.0));
    default:
        return x;
    }
}

static double f_scale(cli_sc k, double x)
{
    if (k == SC_POW10) return pow(10.0, x);
    if (k == SC_POW10_NEG) return pow(10.0, -x);
    return x;
}

/* duration conversion factor from u_par to u_data (transform.c:185-257) */
static double u_duration(const char *u_par, const char *u_data)
{
    static const char *U[] = {"D", "W", "B", "M", "Y"};
    static const double T[5][5] = {
        {1.0, 1.0 / 7.0, 1.0 / 14.0, 12.0 / 365.0, 1.0 / 365.0},
        {7.0, 1.0, 0.5, 84.0 / 365.0, 7.0 / 365.0},
        {14.0, 2.0, 1.0, 84.0 / 365.0, 7.0 / 365.0},
        {365.0 / 12.0, 365.0 / 84.0, 365.0 / 168.0, 1.0, 1.0 / 12.0},
        {365.0, 365.0 / 7.0, 365.0 / 14.0, 12.0, 1.0}};
    int i = -1, j = -1;
    for (int k = 0; k < 5; k++) {
        if (!strcmp(u_par, U[k])) i = k;
        if (!strcmp(u_data, U[k])) j = k;
    }
    if (i < 0 || j < 0) cli_die("invalid units");
    return T[i][j];
}

/* ------------------------------------------------------------------ routers */
static int index_of(const pj *arr, const char *s)
{
    for (int i = 0; i < pj_size(arr); i++)
        if (pj_string(pj_at(arr, i)) && s && !strcmp(pj_string(pj_at(arr, i)), s)) return i;
    return -1;
}

/* position of the series with "id" == id in a par_fixed_values list (-1 if absent) */
static int index_of_series(const pj *series, const char *id)
{
    for (int i = 0; i < pj_size(series); i++) {
        const char *s = pj_string(pj_get(pj_at(series, i), "id"));
        if (s && id && !strcmp(s, id)) return i;
    }
    return -1;
}

static double group_value(const pj *par, const pj *grp, const char *field, double dflt)
{
    const pj *v = pj_get(grp, field);
    if (!v) v = pj_get(par, field); /* unadapted form: value at the parameter level */
    if (pj_is_number(v)) return v->num;
    const pj *vv = pj_get(v, "value");
    if (pj_is_number(vv)) return vv->num;
    return dflt;
}

static const char *group_string(const pj *par, const pj *grp, const char *field)
{
    const pj *v = pj_get(grp, field);
    if (!v) v = pj_get(par, field);
    if (pj_string(v)) return pj_string(v);
    return pj_string(pj_get(v, "value"));
}

/* groups of a partition; the built-in partitions are synthesised when theta lacks them */
static pj *partition_groups(const pj *theta, const char *pid, const pj *order, const char *link_key)
{
    pj *parts = pj_get(theta, "partition");
    pj *p = pj_get(parts, pid);
    if (p) return pj_get(p, "group");
    char text[1 << 16];
    int o = 0;
    if (!strcmp(pid, "variable_population") || !strcmp(pid, "variable_time_series")) {
        o += snprintf(text + o, sizeof text - (size_t)o, "[");
        for (int i = 0; i < pj_size(order); i++)
            o += snprintf(text + o, sizeof text - (size_t)o, "%s{\"id\": \"%s\", \"%s\": [\"%s\"]}", i ? ", " : "",
                          pj_string(pj_at(order, i)), link_key, pj_string(pj_at(order, i)));
        snprintf(text + o, sizeof text - (size_t)o, "]");
    } else if (!strcmp(pid, "identical_population") || !strcmp(pid, "identical_time_series")) {
        o += snprintf(text + o, sizeof text - (size_t)o, "[{\"id\": \"all\", \"%s\": [", link_key);
        for (int i = 0; i < pj_size(order); i++)
            o += snprintf(text + o, sizeof text - (size_t)o, "%s\"%s\"", i ? ", " : "", pj_string(pj_at(order, i)));
        snprintf(text + o, sizeof text - (size_t)o, "]}]");
    } else {
        return NULL;
    }
    char err[256];
    return pj_parse(text, err, sizeof err);   /* synthesised groups live until the command exits */
}

static void build_router(cli_router *r, const char *key, pj *theta, const pj *order, const char *link_key,
                         int is_remainder, int is_bayesian)
{
    memset(r, 0, sizeof *r);
    snprintf(r->name, sizeof r->name, "%s", key);
    r->p = pj_size(order);
    r->map = calloc((size_t)r->p, sizeof(int));
    r->is_remainder = is_remainder;
    r->follow = -1;
    pj *par = NULL;
    const char *pid = "variable_population";
    if (!is_remainder) {
        par = pj_get(pj_get(theta, "parameter"), key);
        const char *fol = pj_string(pj_get(par, "follow"));
        if (fol) par = pj_get(pj_get(theta, "parameter"), fol);
        if (!par) {
            char msg[256];
            snprintf(msg, sizeof msg, "parameter %s follows %s, which is not in theta", key, fol);
            cli_die(msg);
        }
        pid = pj_string(pj_get(par, "partition_id"));
        if (!pid) {
            char msg[256];
            snprintf(msg, sizeof msg, "parameter %s has no partition_id", key);
            cli_die(msg);
        }
    }
    r->par = par;
    pj *groups = partition_groups(theta, pid, order, link_key);
    if (!groups) {
        char msg[256];
        snprintf(msg, sizeof msg, "unknown partition '%s' for %s", pid, key);
        cli_die(msg);
    }
    r->n_gp = pj_size(groups);
    r->group_name = calloc((size_t)r->n_gp, sizeof(char *));
    r->min = calloc((size_t)r->n_gp, sizeof(double));
    r->max = calloc((size_t)r->n_gp, sizeof(double));
    r->guess = calloc((size_t)r->n_gp, sizeof(double));
    r->sd = calloc((size_t)r->n_gp, sizeof(double));
    r->tr = calloc((size_t)r->n_gp, sizeof(cli_tr));
    r->prior = calloc((size_t)r->n_gp, sizeof(cli_prior));
    const char *trs = is_remainder ? "logit" : pj_string(pj_get(par, "transformation"));
    if (!trs) trs = "identity";
    cli_tr tr;
    r->sc = SC_ID;
    if (!strcmp(trs, "log")) tr = TR_LOG;
    else if (!strcmp(trs, "logit")) tr = TR_LOGIT;
    else if (!strcmp(trs, "logit_ab")) tr = TR_LOGIT_AB;
    else if (!strcmp(trs, "identity")) tr = TR_ID;
    else if (!strcmp(trs, "scale_pow10")) { tr = TR_ID; r->sc = SC_POW10; }
    else if (!strcmp(trs, "scale_pow10_bounded")) { tr = TR_LOGIT_AB; r->sc = SC_POW10; }
    else if (!strcmp(trs, "scale_pow10_neg")) { tr = TR_LOG; r->sc = SC_POW10_NEG; }
    else { cli_die("error transf != log, logit, logit_ab, scale_pow10"); return; }
    /* unit multiplier and duration (transform.c:260-330), data unit = days */
    r->multiplier = 1.0;
    const char *unit = pj_string(pj_get(par, "unit"));
    const char *type = pj_string(pj_get(par, "type"));
    if (unit) {
        if (type) {
            if (strcmp(type, "rate_as_duration")) cli_die("not a valid type");
            r->multiplier = u_duration(unit, "D");
            r->is_duration = 1;
        } else {
            r->multiplier = u_duration("D", unit);
        }
    }
    for (int g = 0; g < r->n_gp; g++) {
        pj *grp = pj_at(groups, g);
        const char *gid = pj_string(pj_get(grp, "id"));
        r->group_name[g] = strdup(gid ? gid : "all");
        pj *pg = is_remainder ? NULL : pj_get(pj_get(par, "group"), r->group_name[g]);
        r->min[g] = is_remainder ? 0.0 : group_value(par, pg, "min", 0.0);
        r->max[g] = is_remainder ? 1.0 : group_value(par, pg, "max", 0.0);
        r->guess[g] = is_remainder ? 0.0 : group_value(par, pg, "guess", NAN);
        r->sd[g] = is_remainder ? 0.0 : group_value(par, pg, "sd_transf", 0.0);
        if (!is_remainder && isnan(r->guess[g])) {
            char msg[256];
            snprintf(msg, sizeof msg, "no guess for %s:%s", key, r->group_name[g]);
            cli_die(msg);
        }
        const char *prior = is_remainder ? NULL : group_string(par, pg, "prior");
        r->prior[g] = PRIOR_FLAT;
        if (prior && !strcmp(prior, "normal")) r->prior[g] = PRIOR_NORMAL;
        else if (prior && !strcmp(prior, "pseudo_uniform")) r->prior[g] = PRIOR_PSEUDO_UNIFORM;
        /* Bayesian methods move uniform-prior parameters on the logit scale of [min, max] */
        r->tr[g] = tr;
        if (is_bayesian && prior && (!strcmp(prior, "uniform") || !strcmp(prior, "pseudo_uniform"))) r->tr[g] = TR_LOGIT_AB;
        pj *compo = pj_get(grp, link_key);
        for (int k = 0; k < pj_size(compo); k++) {
            int idx = index_of(order, pj_string(pj_at(compo, k)));
            if (idx < 0) {
                char msg[256];
                snprintf(msg, sizeof msg, "could not find element '%s' in array: '%s'", pj_string(pj_at(compo, k)), link_key);
                cli_die(msg);
            }
            r->map[idx] = g;
        }
    }
}

/* ------------------------------------------------------------------ priors */
static double flat_pdf(double x, double a, double b)
{
    return (x < b && x >= a) ? 1.0 / (b - a) : 0.0;
}

double cli_gauss_pdf(double u, double sd)
{
    return exp(-u * u / (2.0 * sd * sd)) / (sqrt(2.0 * M_PI) * sd);
}

/* normal_prior (metropolis_hastings_prior.c:95-100): mean (min+max)/2, sd (max-min)/4 */
static double normal_prior(double x, double a, double b)
{
    return cli_gauss_pdf(x - 0.5 * (b + a), (b - a) / 4.0);
}

/* pseudo_unif_prior (metropolis_hastings_prior.c:111-132): flat on [min+d, max-d],
 * quadratic shoulders of width d = 0.01 (max-min), integrating to 1 */
static double pseudo_unif_prior(double x, double a, double b)
{
    const double d = 0.01 * (b - a);
    const double h = 1.0 / (b - a - 2.0 / 3.0 * d);
    const double c2 = -h / (d * d);
    const double c1 = -2.0 * (a + d) * c2;
    const double c0 = h + c2 * (a + d) * (a + d);
    if (x <= a || x >= b) return 0.0;
    if (x < a + d) return (c2 * x + c1) * x + c0;
    if (x < b - d) return h;
    const double y = x - (b - a - 2.0 * d);
    return (c2 * y + c1) * y + c0;
}

double cli_prior_pdf(cli_prior kind, double x, double a, double b)
{
    if (kind == PRIOR_NORMAL) return normal_prior(x, a, b);
    if (kind == PRIOR_PSEUDO_UNIFORM) return pseudo_unif_prior(x, a, b);
    return flat_pdf(x, a, b);
}

/* log_prob_prior (metropolis_hastings_prior.c:59-92) over the parameters with sd_transf > 0
 * (neither followers nor the theta remainder, load_best): sum of log prior(f_inv(theta)),
 * each density floored at like_min (NaN, infinite or negative -> like_min and 1 returned) */
int cli_log_prior(const cli_problem *P, const double *theta, double like_min, double *out)
{
    int err = 0;
    double Lp = 0.0;
    for (int i = 0; i < P->n_par; i++) {
        const cli_router *r = &P->R[i];
        if (r->is_remainder || r->follow >= 0) continue;
        for (int g = 0; g < r->n_gp; g++) {
            if (!(r->sd[g] > 0.0)) continue;
            const double x = cli_f_inv(r->tr[g], theta[P->offset[i] + g], r->min[g], r->max[g]);
            double p = cli_prior_pdf(r->prior[g], x, r->min[g], r->max[g]);
            if (isnan(p) || isinf(p) || p < 0.0) {
                err = 1;
                p = like_min;
            } else if (p <= like_min) {
                p = like_min;
            }
            Lp += log(p);
        }
    }
    *out = Lp;
    return err;
}

/* Cholesky factor (lower, row-major) of the n x n matrix s A; returns 0 when positive definite */
int cli_cholesky(const double *A, int n, double s, double *L)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j <= i; j++) {
            double v = s * A[i * n + j];
            for (int k = 0; k < j; k++) v -= L[i * n + k] * L[j * n + k];
            if (i == j) {
                if (!(v > 0.0)) return 1;
                L[i * n + i] = sqrt(v);
            } else {
                L[i * n + j] = v / L[j * n + j];
            }
        }
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) L[i * n + j] = 0.0;
    return 0;
}

/* load_covariance (load.c:413-470): theta's "covariance" into var [n_theta][n_theta], row 0
 * being the header and the theta remainder padded with 0; 0 (var untouched) without one */
int cli_load_covariance(const cli_problem *P, double *var)
{
    pj *cov = pj_get(P->theta, "covariance");
    if (!cov) return 0;
    const int n = P->n_theta;
    int ri = 1;
    for (int i = 0; i < P->n_par; i++)
        for (int g = 0; g < P->R[i].n_gp; g++) {
            const int o = P->offset[i] + g;
            pj *row = NULL;
            if (i != P->ind_theta_remainder) {
                row = pj_at(cov, ri++);
                if (!row || row->type != PJ_ARRAY) cli_die("covariance row is not an array");
            }
            int ck = 0;
            for (int ii = 0; ii < P->n_par; ii++)
                for (int gg = 0; gg < P->R[ii].n_gp; gg++) {
                    const int oo = P->offset[ii] + gg;
                    if (i != P->ind_theta_remainder && ii != P->ind_theta_remainder) {
                        pj *v = pj_at(row, ck++);
                        if (!pj_is_number(v)) cli_die("covariance entry is not a number");
                        var[o * n + oo] = v->num;
                    } else {
                        var[o * n + oo] = 0.0;
                    }
                }
        }
    return 1;
}

double cli_natural(const cli_problem *P, int i, int g, const double *theta)
{
    const cli_router *r = &P->R[i];
    double x = cli_f_inv(r->tr[g], theta[P->offset[i] + g], r->min[g], r->max[g]);
    x = f_scale(r->sc, x);
    x *= r->multiplier;
    if (r->is_duration) x = 1.0 / (0.00001 + x);
    return x;
}

/* ------------------------------------------------------------------ loading */
double CLI_FREEZE = -1.0;

/* The series of covariate slot s: par_fixed_values[name] entry "id" == key (NULL if absent):
 * x in days since date_0, y times the unit multiplier (get_multiplier, transform.c:267-301,
 * data unit = days); frozen (-g) at its value at min(freeze, last knot) over its knots
 * (build.c:1188-1205). */
static void load_series(cli_problem *P, int s, const pj *list, const char *key)
{
    plom_covariate *c = &P->cov[s];
    const int q = index_of_series(list, key);
    if (q < 0) {
        /* no series for this population / series: the slot is never read */
        c->n = 1;
        c->x = calloc(1, sizeof(double));
        c->y = calloc(1, sizeof(double));
        return;
    }
    const pj *ser = pj_at(list, q), *jx = pj_get(ser, "x"), *jy = pj_get(ser, "y");
    const int n = pj_size(jy);
    if (n < 1 || pj_size(jx) != n) cli_die("a par_fixed_values series needs as many x as y values");
    double *x = calloc((size_t)n, sizeof(double)), *y = calloc((size_t)n, sizeof(double));
    double mult = 1.0;
    const char *unit = pj_string(pj_get(ser, "unit")), *type = pj_string(pj_get(ser, "type"));
    if (unit) {
        if (type && strcmp(type, "rate_as_duration")) cli_die("not a valid type");
        mult = type ? u_duration(unit, "D") : u_duration("D", unit);
    }
    for (int i = 0; i < n; i++) {
        x[i] = pj_number(pj_at(jx, i), NULL);
        y[i] = pj_number(pj_at(jy, i), NULL) * mult;
    }
    c->n = n;
    c->x = x;
    c->y = y;
    /* the value at t = 0 before any freeze: the population size of the initial state
     * (pop_size_t0 is set before the freeze, build.c:1179-1205) */
    const double t0 = 0.0;
    if (plom_covariate_eval(c, 1, &t0, &P->cov_raw0[s]) != PLOM_OK) cli_die(plom_last_error());
    if (CLI_FREEZE >= 0.0 && n > 1) {
        double t = CLI_FREEZE < x[n - 1] ? CLI_FREEZE : x[n - 1], v;
        if (plom_covariate_eval(c, 1, &t, &v) != PLOM_OK) cli_die(plom_last_error());
        double *fx = calloc(2, sizeof(double)), *fy = calloc(2, sizeof(double));
        fx[0] = x[0];
        fx[1] = x[n - 1];
        fy[0] = fy[1] = v;
        c->n = 2;
        c->x = fx;
        c->y = fy;
    }
    for (int i = 1; i < c->n; i++)
        if (c->y[i] != c->y[0]) P->has_cov = 1;
}

/* slots: [covariate][N_CAC] then [observation covariate][N_TS] (an observation covariate
 * without a series of its own for a time series takes its first population's); the values
 * at t = 0 of the (frozen) series go into pv, those of the series as given make the
 * population sizes of the initial state (pop_size_t0, build.c:1179-1185) */
static void cli_load_covariates(cli_problem *P, const pj *pfv, const pj *par_fixed, const pj *pfo)
{
    const int N_CAC = P->N_CAC, N_TS = P->N_TS;
    const int n_fix = P->n_fixed * N_CAC + pj_size(pfo) * N_TS;
    if (n_fix != P->mi.n_fix) cli_die("settings covariates differ from the compiled model");
    P->cov = calloc((size_t)n_fix + 1, sizeof(plom_covariate));
    P->cov0 = calloc((size_t)n_fix + 1, sizeof(double));
    P->cov_raw0 = calloc((size_t)n_fix + 1, sizeof(double));
    int s = 0;
    for (int k = 0; k < P->n_fixed; k++) {
        const pj *list = pj_get(pfv, pj_string(pj_at(par_fixed, k)));
        for (int cac = 0; cac < N_CAC; cac++) load_series(P, s++, list, pj_string(pj_at(P->cac_id, cac)));
    }
    for (int q = 0; q < pj_size(pfo); q++) {
        const pj *list = pj_get(pfv, pj_string(pj_at(pfo, q)));
        int k = index_of(par_fixed, pj_string(pj_at(pfo, q)));
        for (int ts = 0; ts < N_TS; ts++) {
            if (index_of_series(list, pj_string(pj_at(P->ts_id, ts))) >= 0 || k < 0) {
                load_series(P, s++, list, pj_string(pj_at(P->ts_id, ts)));
            } else {
                P->cov[s] = P->cov[k * N_CAC];
                P->cov_raw0[s] = P->cov_raw0[k * N_CAC];
                s++;
            }
        }
    }
    const double t0 = 0.0;
    for (int i = 0; i < n_fix; i++)
        if (plom_covariate_eval(&P->cov[i], 1, &t0, &P->cov0[i]) != PLOM_OK) cli_die(plom_last_error());
    for (int i = 0; i < P->n_fixed * N_CAC; i++) P->fixed[i] = P->cov_raw0[i];
}

void cli_load(cli_problem *P, int is_bayesian, int nb_obs, double dt)
{
    char err[512], str[1024];
    memset(P, 0, sizeof *P);
    P->settings = pj_load_file("./.settings.json", err, sizeof err);
    if (!P->settings) cli_die(err);
    char *line = malloc(BUFSIZE);
    if (!fgets(line, BUFSIZE, stdin)) cli_die("no theta on stdin");
    P->theta = pj_parse(line, err, sizeof err);
    free(line);
    if (!P->theta) cli_die(err);

    pj *settings = P->settings, *theta = P->theta;
    pj *cst = pj_get(settings, "cst"), *orders = pj_get(settings, "orders"), *sdata = pj_get(settings, "data");
    const int N_C = (int)pj_number(pj_get(cst, "N_C"), NULL), N_AC = (int)pj_number(pj_get(cst, "N_AC"), NULL);
    P->N_CAC = N_C * N_AC;
    P->N_TS = (int)pj_number(pj_get(cst, "N_TS"), NULL);
    P->N_DATA = (int)pj_number(pj_get(cst, "N_DATA"), NULL);
    P->model_name = pj_string(pj_get(settings, "model"));
    if (!P->model_name) cli_die(".settings.json has no 'model' (write it with this build's model builder)");
    /* a model outside the library: the model library plom_build.py --lib wrote beside the
     * settings (a path relative to the current directory, where .settings.json is read) */
    const char *mlib = pj_string(pj_get(settings, "model_library"));
    if (mlib) {
        int32_t mid = -1;
        if (plom_model_load(mlib, &mid) != PLOM_OK) cli_die(plom_last_error());
    }
    P->model_id = plom_model_find(P->model_name);
    if (P->model_id < 0) cli_die(plom_last_error());
    plom_model_info_get(P->model_id, &P->mi);
    if (P->mi.n_cac != P->N_CAC || P->mi.n_ts != P->N_TS) cli_die("settings dimensions differ from the compiled model");
    const int N_CAC = P->N_CAC, N_TS = P->N_TS;

    P->par_sv = pj_get(orders, "par_sv");
    pj *par_proc = pj_get(orders, "par_proc"), *par_obs = pj_get(orders, "par_obs");
    pj *par_fixed = pj_get(orders, "par_fixed");
    P->cac_id = pj_get(orders, "cac_id");
    P->ts_id = pj_get(orders, "ts_id");
    P->n_sv = pj_size(P->par_sv);
    P->n_proc = pj_size(par_proc);
    P->n_obs = pj_size(par_obs);
    P->n_par = P->n_sv + P->n_proc + P->n_obs;
    P->R = calloc((size_t)P->n_par, sizeof(cli_router));
    P->offset = calloc((size_t)P->n_par + 1, sizeof(int));
    P->ind_theta_remainder = -1;
    for (int i = 0; i < P->n_par; i++) {
        const pj *lst = i < P->n_sv ? P->par_sv : (i < P->n_sv + P->n_proc ? par_proc : par_obs);
        int k = i < P->n_sv ? i : (i < P->n_sv + P->n_proc ? i - P->n_sv : i - P->n_sv - P->n_proc);
        const char *key = pj_string(pj_at(lst, k));
        const int is_obs = i >= P->n_sv + P->n_proc;
        int missing = pj_get(pj_get(theta, "parameter"), key) == NULL;
        if (missing && i >= P->n_sv) {
            snprintf(str, sizeof str, "parameter %s missing from theta", key);
            cli_die(str);
        }
        if (missing) P->ind_theta_remainder = i;
        build_router(&P->R[i], key, theta, is_obs ? P->ts_id : P->cac_id, is_obs ? "time_series_id" : "population_id",
                     missing, is_bayesian);
        P->offset[i + 1] = P->offset[i] + P->R[i].n_gp;
    }
    P->n_theta = P->offset[P->n_par];
    /* followers (build.c:1540-1589, build.c:425-441): same kind of parameter as the followed one */
    for (int i = 0; i < P->n_par; i++) {
        if (P->R[i].is_remainder) continue;
        const char *fol = pj_string(pj_get(pj_get(pj_get(theta, "parameter"), P->R[i].name), "follow"));
        if (!fol) continue;
        int j = -1;
        for (int q = 0; q < P->n_par; q++)
            if (!strcmp(P->R[q].name, fol)) j = q;
        const int obs_i = i >= P->n_sv + P->n_proc, obs_j = j >= P->n_sv + P->n_proc;
        if (j < 0 || obs_i != obs_j || P->R[j].n_gp != P->R[i].n_gp) {
            snprintf(str, sizeof str, "invalid parameter type between %s (follower) and %s (follow)", P->R[i].name, fol);
            cli_die(str);
        }
        P->R[i].follow = j;
    }
    /* drifting parameters: the groups and transforms of the parameter's router; the
     * volatility of group g is natural[vol][g] (drift.c:38-40), i.e. the pv slot of the first
     * population in group g of the volatility's router */
    pj *jdr = pj_get(settings, "drift");
    pj *jdp = pj_get(jdr, "ind_par_Xdrift_applied"), *jdv = pj_get(jdr, "ind_volatility_Xdrift");
    P->n_drift = pj_size(jdp);
    if (P->n_drift != P->mi.n_drift || pj_size(jdv) != P->n_drift) cli_die("settings drift differs from the compiled model");
    if (P->n_drift) {
        const size_t nd = (size_t)P->n_drift * N_CAC;
        int32_t *grp = calloc(nd, 4), *tr = calloc(nd, 4), *sc = calloc(nd, 4), *dur = calloc(nd, 4), *vs = calloc(nd, 4);
        double *mn = calloc(nd, 8), *mx = calloc(nd, 8), *mult = calloc(nd, 8);
        P->drift_par = calloc((size_t)P->n_drift, sizeof(int));
        for (int i = 0; i < P->n_drift; i++) {
            const int ip = (int)pj_number(pj_at(jdp, i), NULL), iv = (int)pj_number(pj_at(jdv, i), NULL);
            if (ip < P->n_sv || ip >= P->n_sv + P->n_proc || iv < P->n_sv || iv >= P->n_sv + P->n_proc)
                cli_die("a drifting parameter and its volatility must be process parameters");
            P->drift_par[i] = ip;
            const cli_router *r = &P->R[ip], *rv = &P->R[iv];
            for (int c = 0; c < N_CAC; c++) {
                const int g = r->map[c], q = i * N_CAC + c;
                grp[q] = g;
                tr[q] = (int32_t)r->tr[g];
                sc[q] = (int32_t)r->sc;
                dur[q] = r->is_duration;
                mn[q] = r->min[g];
                mx[q] = r->max[g];
                mult[q] = r->multiplier;
                vs[q] = -1;
            }
            for (int g = 0; g < N_CAC; g++)
                for (int c = 0; c < N_CAC; c++)
                    if (rv->map[c] == g) {
                        vs[i * N_CAC + g] = (iv - P->n_sv) * N_CAC + c;
                        break;
                    }
        }
        P->dmap.group = grp; P->dmap.tr = tr; P->dmap.scale = sc; P->dmap.duration = dur;
        P->dmap.min = mn; P->dmap.max = mx; P->dmap.multiplier = mult; P->dmap.vol_slot = vs;
    }
    P->noise_par = calloc((size_t)P->n_par, sizeof(int));
    pj *ins = pj_get(settings, "ind_noise_sd");
    for (int q = 0; q < pj_size(ins); q++) {
        int k = (int)pj_number(pj_at(ins, q), NULL);
        if (k >= 0 && k < P->n_par) P->noise_par[k] = 1;
    }

    /* covariates: one series per covariate slot of pv (build_p_calc, build.c:1075-1211) */
    P->n_fixed = pj_size(par_fixed);
    P->fixed = calloc((size_t)P->n_fixed * N_CAC + 1, sizeof(double));
    cli_load_covariates(P, pj_get(sdata, "par_fixed_values"), par_fixed, pj_get(settings, "par_fixed_obs"));

    /* data and times (build.c:662-881) */
    P->N = (nb_obs > 0 && nb_obs < P->N_DATA) ? nb_obs : P->N_DATA;
    const int N = P->N;
    P->times = calloc((size_t)N + 1, sizeof(double));
    P->data = calloc((size_t)N * N_TS + 1, sizeof(double));
    pj *jt = pj_get(sdata, "times"), *jd = pj_get(sdata, "data");
    for (int n = 0; n <= N; n++) P->times[n] = pj_number(pj_at(jt, n), NULL);
    for (int n = 0; n < N; n++)
        for (int ts = 0; ts < N_TS; ts++) {
            pj *v = pj_at(pj_at(jd, n), ts);
            P->data[n * N_TS + ts] = pj_is_number(v) ? v->num : NAN;
        }
    if (dt <= 0.0) dt = 0.25;
    P->dt = 1.0 / ((double)round(1.0 / dt));

    P->remainder = pj_string(pj_get(settings, "remainder"));
    const pj *psz = pj_get(settings, "POP_SIZE_EQ_SUM_SV");
    P->has_rem = !(psz && psz->type == PJ_TRUE) && P->remainder;
    /* the population sizes at the observation times (remainders of the X file) */
    P->pop_t = calloc((size_t)(N + 1) * N_CAC + 1, sizeof(double));
    double *col = calloc((size_t)N + 2, sizeof(double));
    for (int cac = 0; P->n_fixed > 0 && cac < N_CAC; cac++) {
        if (plom_covariate_eval(&P->cov[cac], N + 1, P->times, col) != PLOM_OK) cli_die(plom_last_error());
        for (int n = 0; n <= N; n++) P->pop_t[n * N_CAC + cac] = col[n];
    }
    free(col);
}

void cli_theta_from_guess(const cli_problem *P, double *theta)
{
    for (int i = 0; i < P->n_par; i++)
        for (int g = 0; g < P->R[i].n_gp; g++)
            theta[P->offset[i] + g] = cli_f_tr(P->R[i].tr[g], P->R[i].guess[g], P->R[i].min[g], P->R[i].max[g]);
    if (cli_theta_remainder(P, theta)) cli_die("sum of the initial conditions without remainder is greater than 1.0");
}

int cli_theta_remainder(const cli_problem *P, double *theta)
{
    int n_err = 0;
    for (int cac = 0; cac < P->N_CAC; cac++) {
        double prop = 0.0;
        for (int i = 0; i < P->n_sv; i++)
            if (i != P->ind_theta_remainder) prop += cli_natural(P, i, P->R[i].map[cac], theta);
        if (prop > 1.0) {
            n_err++;
        } else if (P->ind_theta_remainder >= 0) {
            const int i = P->ind_theta_remainder;
            theta[P->offset[i] + cac] = cli_f_tr(TR_LOGIT, 1.0 - prop, 0.0, 1.0);
        }
    }
    return n_err;
}

void cli_follow(const cli_problem *P, double *theta)
{
    for (int i = 0; i < P->n_par; i++)
        if (P->R[i].follow >= 0)
            for (int g = 0; g < P->R[i].n_gp; g++) theta[P->offset[i] + g] = theta[P->offset[P->R[i].follow] + g];
}

void cli_pack(const cli_problem *P, const double *theta, double *pv, double *x0)
{
    const int N_CAC = P->N_CAC, N_TS = P->N_TS;
    char str[256];
    /* parameter vector in the generated layout: proc [n_proc][N_CAC], obs [n_obs][N_TS],
     * fixed [n_fixed][N_CAC], fixed_obs [..][N_TS], coupling [N_CAC][N_CAC] */
    int o = 0;
    for (int i = P->n_sv; i < P->n_sv + P->n_proc; i++)
        for (int cac = 0; cac < N_CAC; cac++) pv[o++] = cli_natural(P, i, P->R[i].map[cac], theta);
    for (int i = P->n_sv + P->n_proc; i < P->n_par; i++)
        for (int ts = 0; ts < N_TS; ts++) pv[o++] = cli_natural(P, i, P->R[i].map[ts], theta);
    for (int k = 0; k < P->mi.n_fix; k++) pv[o++] = P->cov0[k];   /* covariates at t = 0 */
    pj *coupling = pj_get(pj_get(P->settings, "coupling"), "matrix");
    for (int a = 0; a < pj_size(coupling); a++)
        for (int b = 0; b < pj_size(pj_at(coupling, a)); b++) pv[o++] = pj_number(pj_at(pj_at(coupling, a), b), NULL);
    /* school terms (IS_SCHOOL_TERMS, build.c:834-868): per population n_terms, prop_school =
     * 1 - sum(end - begin), then the (begin, end) pairs of data.school_terms, in years */
    if (pj_number(pj_get(pj_get(P->settings, "cst"), "IS_SCHOOL_TERMS"), NULL) == 1.0) {
        pj *terms = pj_get(pj_get(P->settings, "data"), "school_terms");
        if (pj_size(terms) != N_CAC) cli_die("settings data.school_terms needs one list per population");
        for (int cac = 0; cac < N_CAC; cac++) {
            pj *lst = pj_at(terms, cac);
            const int n = pj_size(lst) / 2;
            if (pj_size(lst) % 2 || n > PLOM_MAX_TERMS) cli_die("settings data.school_terms: (begin, end) pairs, at most 8 per population");
            double *rec = pv + o;
            memset(rec, 0, PLOM_SCHOOL_REC * sizeof(double));
            rec[0] = n;
            double held = 0.0;
            for (int k = 0; k < 2 * n; k++) rec[2 + k] = pj_number(pj_at(lst, k), NULL);
            for (int k = 0; k < n; k++) held += rec[3 + 2 * k] - rec[2 + 2 * k];
            rec[1] = 1.0 - held;
            o += PLOM_SCHOOL_REC;
        }
    }
    if (o != P->mi.pv_size) {
        snprintf(str, sizeof str, "parameter vector has %d values, the model expects %d", o, P->mi.pv_size);
        cli_die(str);
    }
    /* initial state: proportions x population size (prediction_util.c:82-115) */
    memset(x0, 0, (size_t)P->mi.size_proj * sizeof(double));
    for (int i = 0; i < P->n_sv; i++)
        for (int cac = 0; cac < N_CAC; cac++) {
            x0[i * N_CAC + cac] = cli_natural(P, i, P->R[i].map[cac], theta) * P->fixed[cac];
            if (P->psr) x0[i * N_CAC + cac] = round(x0[i * N_CAC + cac]);   /* counts (prediction_util.c:108-110) */
        }
    /* drift states: the drifting parameter's theta (theta_driftIC2Xdrift, prediction_util.c:119-131) */
    for (int i = 0; i < P->n_drift; i++)
        for (int cac = 0; cac < N_CAC; cac++) {
            const int ip = P->drift_par[i];
            x0[P->mi.drift_offset + i * N_CAC + cac] = theta[P->offset[ip] + P->R[ip].map[cac]];
        }
}

/* ------------------------------------------------------------------ writers */
FILE *cli_fopen(const char *path, const char *name, int id)
{
    char str[2048];
    snprintf(str, sizeof str, "%s%s_%d.csv", path, name, id);
    FILE *f = fopen(str, "w");
    if (!f) {
        char msg[2100];
        snprintf(msg, sizeof msg, "cannot write %s", str);
        cli_die(msg);
    }
    return f;
}

void cli_trace_header(const cli_problem *P, FILE *f)
{
    fprintf(f, "index,");
    for (int i = 0; i < P->n_par; i++)
        if (i != P->ind_theta_remainder)
            for (int g = 0; g < P->R[i].n_gp; g++) fprintf(f, "%s:%s,", P->R[i].name, P->R[i].group_name[g]);
    fprintf(f, "log_like\n");
}

void cli_trace_row(const cli_problem *P, FILE *f, int index, const double *theta, double log_like)
{
    const char *fmt = CLI_TRACE_17G ? "%.17g," : "%g,";
    fprintf(f, "%d,", index);
    for (int i = 0; i < P->n_par; i++)
        if (i != P->ind_theta_remainder)
            for (int g = 0; g < P->R[i].n_gp; g++)
                fprintf(f, fmt, cli_f_inv(P->R[i].tr[g], theta[P->offset[i] + g], P->R[i].min[g], P->R[i].max[g]));
    fprintf(f, CLI_TRACE_17G ? "%.17g\n" : "%g\n", log_like);
}

void cli_x_header(const cli_problem *P, FILE *f)
{
    fprintf(f, "index,time,");
    for (int i = 0; i < P->n_sv; i++)
        for (int cac = 0; cac < P->N_CAC; cac++)
            fprintf(f, "%s:%s,", pj_string(pj_at(P->par_sv, i)), pj_string(pj_at(P->cac_id, cac)));
    if (P->has_rem)
        for (int cac = 0; cac < P->N_CAC; cac++) fprintf(f, "%s:%s,", P->remainder, pj_string(pj_at(P->cac_id, cac)));
    for (int ts = 0; ts < P->N_TS; ts++) fprintf(f, "obs_mean:%s,", pj_string(pj_at(P->ts_id, ts)));
    /* drift columns (header_X, print.c:115-122) */
    for (int i = 0; P->drift_on && i < P->n_drift; i++) {
        const cli_router *r = &P->R[P->drift_par[i]];
        for (int g = 0; g < r->n_gp; g++) fprintf(f, "drift:%s:%s,", r->name, r->group_name[g]);
    }
    for (int ts = 0; ts < P->N_TS; ts++)
        fprintf(f, "obs_real:%s%s", pj_string(pj_at(P->ts_id, ts)), ts < P->N_TS - 1 ? "," : "");
    fprintf(f, "\n");
}

void cli_x_rows(const cli_problem *P, FILE *f, int index, const double *lineage, const double *lineage_obs)
{
    const int S = P->mi.size_proj, N_TS = P->N_TS;
    for (int r = P->N; r >= 0; r--) cli_x_row(P, f, index, r, lineage + (size_t)r * S, lineage_obs + (size_t)r * N_TS * 2);
}

void cli_x_row(const cli_problem *P, FILE *f, int index, int r, const double *x, const double *ob)
{
    const int N_CAC = P->N_CAC, N_TS = P->N_TS;
    {
        fprintf(f, "%d,%g,", index, P->times[r]);
        for (int k = 0; k < P->n_sv * N_CAC; k++) fprintf(f, "%g,", x[k]);
        if (P->has_rem)
            for (int cac = 0; cac < N_CAC; cac++) {
                double s = 0.0;
                for (int i = 0; i < P->n_sv; i++) s += x[i * N_CAC + cac];
                fprintf(f, "%g,", P->pop_t[r * N_CAC + cac] - s);
            }
        for (int ts = 0; ts < N_TS; ts++) fprintf(f, "%g,", ob[2 * ts]);
        /* drift: f_inv of the state of the group's first population (print_X, print.c:321-327) */
        for (int i = 0; P->drift_on && i < P->n_drift; i++) {
            const cli_router *rr = &P->R[P->drift_par[i]];
            for (int g = 0; g < rr->n_gp; g++)
                for (int c = 0; c < N_CAC; c++)
                    if (rr->map[c] == g) {
                        fprintf(f, "%g,", cli_f_inv(rr->tr[g], x[P->mi.drift_offset + i * N_CAC + c], rr->min[g], rr->max[g]));
                        break;
                    }
        }
        for (int ts = 0; ts < N_TS; ts++) fprintf(f, "%g%s", ob[2 * ts + 1], ts < N_TS - 1 ? "," : "");
        fprintf(f, "\n");
    }
}

void cli_done(const cli_problem *P, const double *theta, const char *path, int id, int pipe)
{
    for (int i = 0; i < P->n_par; i++) {
        const cli_router *r = &P->R[i];
        if (i == P->ind_theta_remainder || !r->par) continue;
        pj *par = pj_get(pj_get(P->theta, "parameter"), r->name);
        pj *groups = pj_get(par, "group");
        for (int g = 0; g < r->n_gp; g++) {
            pj *grp = pj_get(groups, r->group_name[g]);
            pj *guess = pj_get(grp, "guess");
            if (guess && guess->type == PJ_OBJECT)
                pj_set(guess, "value", pj_new_real(cli_f_inv(r->tr[g], theta[P->offset[i] + g], r->min[g], r->max[g])));
        }
    }
    if (pipe) {
        pj_dump(P->theta, stdout);
        fflush(stdout);
    } else {
        char str[2048];
        snprintf(str, sizeof str, "%s.done_%d.json", path, id);
        FILE *fd = fopen(str, "w");
        if (fd) {
            pj_dump(P->theta, fd);
            fclose(fd);
        }
    }
}
