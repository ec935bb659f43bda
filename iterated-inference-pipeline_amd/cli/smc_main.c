/*
 * smc_main.c -- the drop-in `smc` command: same options, inputs and outputs as the
 * reference's model_builder/C/smc/main.c, with the filter running on the GPU through
 * the C ABI (include/plom_smc.h).
 *
 *   cat theta.json | ./smc sde -J 100000 -P | ...
 *
 * Inputs: ./.settings.json (written by this build's model builder in the schema of
 * model_builder/Builder.py:173-259, plus the keys "model", "par_fixed_obs" and
 * "coupling" naming the generated model compiled into the library) and one line of
 * theta JSON on stdin (json.c:359-383).
 * Outputs: the log lines of main.c:266-356 (unless -q/-P), trace_<id>.csv
 * (print.c:189-207, 396-440), hat_<id>.csv (print.c:151-186, 445-573: time, low95 / mean
 * / high95 of every state, remainder and observation mean, rows t = 0 and every
 * interval), pred_res_<id>.csv (print.c:132-143, 632-701: rows for the intervals with
 * data), and the theta with updated guesses on stdout (-P) or in .done_<id>.json
 * (print.c:33-67).
 *
 * Parameter handling follows the reference: routers from partitions
 * (build.c:267-475), transform then back-transform with scale, unit multiplier and
 * duration->rate (transform.c:23-107, 185-301, 462-514), the theta remainder
 * (json.c:405-424, proposal.c:87-114), initial state = proportions x population size at
 * t = 0 (prediction_util.c:82-115, build.c:1179-1185).
 *
 * Differences: implementation "psr" and "ode" are not built (only "sde"); the X output
 * of -r is not written (a note goes to stderr); hat and pred_res use this build's
 * arithmetic (exact integer weights for the 95 % bounds, tree sums);
 * --prior is not supported; -N is accepted and ignored (the GPU replaces threads).
 * Extensions: --fp32 (FP32 propagation), --device <k>, --seed <k> (random key, default
 * 2 + id as build.c:965-971), a full-precision "logV_17g" log line, and --dry_run (print
 * the natural parameter vector and initial state as JSON and stop before the GPU).
 */
#include <getopt.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "plom_smc.h"
#include "pjson.h"

#define ZERO_LOG 1e-17
#define ONE_LOGIT 0.999999999
#define BUFSIZE (5000 * 1024)

static int QUIET = 0;

static void print_log(const char *msg)
{
    if (!QUIET) printf("%s\n", msg);
}

static void die(const char *msg)
{
    fprintf(stderr, "error: %s\n", msg);
    exit(EXIT_FAILURE);
}

/* ------------------------------------------------------------------ transforms */
typedef enum { TR_ID, TR_LOG, TR_LOGIT, TR_LOGIT_AB } tr_kind;
typedef enum { SC_ID, SC_POW10, SC_POW10_NEG } sc_kind;

static double f_tr(tr_kind k, double x, double a, double b)
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

static double f_inv(tr_kind k, double x, double a, double b)
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

static double f_scale(sc_kind k, double x)
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
    if (i < 0 || j < 0) die("invalid units");
    return T[i][j];
}

/* ------------------------------------------------------------------ routers */
typedef struct {
    char name[128];
    int p;              /* number of populations (or series) */
    int n_gp;
    int *map;           /* element -> group */
    char **group_name;
    double *min, *max, *guess, *mean;
    tr_kind tr;
    sc_kind sc;
    double multiplier;
    int is_duration;
    int is_remainder;
    pj *par;            /* the theta parameter object (NULL for the remainder) */
} router;

static int index_of(const pj *arr, const char *s)
{
    for (int i = 0; i < pj_size(arr); i++)
        if (pj_string(pj_at(arr, i)) && !strcmp(pj_string(pj_at(arr, i)), s)) return i;
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

/* groups of a partition; the built-in partitions are synthesised when theta lacks them */
static pj *partition_groups(const pj *theta, const char *pid, const pj *order, const char *link_key)
{
    pj *parts = pj_get(theta, "partition");
    pj *p = pj_get(parts, pid);
    if (p) return pj_get(p, "group");
    pj *groups = NULL;   /* synthesised groups live until the command exits */
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
    groups = pj_parse(text, err, sizeof err);
    return groups;
}

static void build_router(router *r, const char *key, pj *theta, const pj *order, const char *link_key, int is_remainder)
{
    memset(r, 0, sizeof *r);
    snprintf(r->name, sizeof r->name, "%s", key);
    r->p = pj_size(order);
    r->map = calloc((size_t)r->p, sizeof(int));
    r->is_remainder = is_remainder;
    pj *par = NULL;
    const char *pid = "variable_population";
    if (!is_remainder) {
        par = pj_get(pj_get(theta, "parameter"), key);
        const char *fol = pj_string(pj_get(par, "follow"));
        if (fol) par = pj_get(pj_get(theta, "parameter"), fol);
        pid = pj_string(pj_get(par, "partition_id"));
        if (!pid) {
            char msg[256];
            snprintf(msg, sizeof msg, "parameter %s has no partition_id", key);
            die(msg);
        }
    }
    r->par = par;
    pj *groups = partition_groups(theta, pid, order, link_key);
    if (!groups) {
        char msg[256];
        snprintf(msg, sizeof msg, "unknown partition '%s' for %s", pid, key);
        die(msg);
    }
    r->n_gp = pj_size(groups);
    r->group_name = calloc((size_t)r->n_gp, sizeof(char *));
    r->min = calloc((size_t)r->n_gp, sizeof(double));
    r->max = calloc((size_t)r->n_gp, sizeof(double));
    r->guess = calloc((size_t)r->n_gp, sizeof(double));
    r->mean = calloc((size_t)r->n_gp, sizeof(double));
    const char *tr = is_remainder ? "logit" : pj_string(pj_get(par, "transformation"));
    if (!tr) tr = "identity";
    r->sc = SC_ID;
    if (!strcmp(tr, "log")) r->tr = TR_LOG;
    else if (!strcmp(tr, "logit")) r->tr = TR_LOGIT;
    else if (!strcmp(tr, "logit_ab")) r->tr = TR_LOGIT_AB;
    else if (!strcmp(tr, "identity")) r->tr = TR_ID;
    else if (!strcmp(tr, "scale_pow10")) { r->tr = TR_ID; r->sc = SC_POW10; }
    else if (!strcmp(tr, "scale_pow10_bounded")) { r->tr = TR_LOGIT_AB; r->sc = SC_POW10; }
    else if (!strcmp(tr, "scale_pow10_neg")) { r->tr = TR_LOG; r->sc = SC_POW10_NEG; }
    else die("error transf != log, logit, logit_ab, scale_pow10");
    /* unit multiplier and duration (transform.c:260-330), data unit = days */
    r->multiplier = 1.0;
    const char *unit = pj_string(pj_get(par, "unit"));
    const char *type = pj_string(pj_get(par, "type"));
    if (unit) {
        if (type) {
            if (strcmp(type, "rate_as_duration")) die("not a valid type");
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
        if (!is_remainder && isnan(r->guess[g])) {
            char msg[256];
            snprintf(msg, sizeof msg, "no guess for %s:%s", key, r->group_name[g]);
            die(msg);
        }
        pj *compo = pj_get(grp, link_key);
        for (int k = 0; k < pj_size(compo); k++) {
            int idx = index_of(order, pj_string(pj_at(compo, k)));
            if (idx < 0) {
                char msg[256];
                snprintf(msg, sizeof msg, "could not find element '%s' in array: '%s'", pj_string(pj_at(compo, k)), link_key);
                die(msg);
            }
            r->map[idx] = g;
        }
    }
}

/* back_transform_x (transform.c:462-480) */
static double natural(const router *r, int g)
{
    double x = f_inv(r->tr, r->mean[g], r->min[g], r->max[g]);
    x = f_scale(r->sc, x);
    x *= r->multiplier;
    if (r->is_duration) x = 1.0 / (0.00001 + x);
    return x;
}

/* ------------------------------------------------------------------ main */
int main(int argc, char *argv[])
{
    int filter = 1, J = 1, nb_obs = -1, general_id = 0, n_threads = 1, fp32 = 0, device = 0;
    int print_best = 1, print_hat = 1, print_pred_res = 1, print_X = 0, pipe = 0, prior = 0, dry_run = 0;
    long seed_opt = -1;
    double dt = 0.0, like_min = 1e-17;
    unsigned int noises_off = 0;
    char path[1024] = "./";
    static struct option long_options[] = {
        {"traj", no_argument, 0, 'r'}, {"no_dem_sto", no_argument, 0, 'x'},
        {"no_white_noise", no_argument, 0, 'y'}, {"no_diff", no_argument, 0, 'z'},
        {"DT", required_argument, 0, 's'}, {"eps_abs", required_argument, 0, 'v'},
        {"eps_rel", required_argument, 0, 'w'}, {"nb_obs", required_argument, 0, 'o'},
        {"freeze_forcing", required_argument, 0, 'g'}, {"help", no_argument, 0, 'h'},
        {"path", required_argument, 0, 'p'}, {"prior", no_argument, 0, 1000},
        {"id", required_argument, 0, 'i'}, {"n_thread", required_argument, 0, 'N'},
        {"no_filter", no_argument, 0, 't'}, {"no_trace", no_argument, 0, 'b'},
        {"no_hat", no_argument, 0, 'e'}, {"no_pred_res", no_argument, 0, 'd'},
        {"quiet", no_argument, 0, 'q'}, {"pipe", no_argument, 0, 'P'},
        {"LIKE_MIN", required_argument, 0, 'l'}, {"fp32", no_argument, 0, 1001},
        {"device", required_argument, 0, 1002}, {"seed", required_argument, 0, 1003},
        {"dry_run", no_argument, 0, 1004}, {0, 0, 0, 0}};
    const char *help =
        "PLOM Sequential Monte Carlo (MI355X)\n"
        "usage:\n"
        "smc [implementation] [--no_dem_sto] [--no_white_noise] [--no_diff]\n"
        "                     [-r, --traj] [-p, --path <path>] [-i, --id <integer>] [-N, --n_thread <integer>]\n"
        "                     [-g, --freeze_forcing <float>]\n"
        "                     [-t, --no_filter] [-b, --no_trace] [-e, --no_hat]\n"
        "                     [-s, --DT <float>] [--eps_abs <float>] [--eps_rel <float>]\n"
        "                     [-l, --LIKE_MIN <float>] [-J <integer>] [--prior]\n"
        "                     [-q, --quiet] [-P, --pipe] [--fp32] [--device <k>] [--seed <k>] [-h, --help]\n"
        "where implementation is 'sde' (this build runs the SDE filter on the GPU)\n";
    for (;;) {
        int idx = 0;
        int ch = getopt_long(argc, argv, "qPxyzs:v:w:p:i:J:l:trbehrN:o:g:d", long_options, &idx);
        if (ch == -1) break;
        switch (ch) {
        case 'x': noises_off |= PLOM_NO_DEM_STO_BIT; break;
        case 'y': noises_off |= PLOM_NO_ENV_STO_BIT; break;
        case 'z': noises_off |= PLOM_NO_DRIFT_BIT; break;
        case 's': dt = atof(optarg); break;
        case 'v': case 'w': case 'g': break;
        case 'o': nb_obs = atoi(optarg); break;
        case 'h': printf("%s", help); return 1;
        case 'p': snprintf(path, sizeof path, "%s", optarg); break;
        case 'i': general_id = atoi(optarg); break;
        case 't': filter = 0; break;
        case 'J': J = atoi(optarg); break;
        case 'N': n_threads = atoi(optarg); break;
        case 'l': like_min = atof(optarg); break;
        case 'r': print_X = 1; break;
        case 'b': print_best = 0; break;
        case 'e': print_hat = 0; break;
        case 'd': print_pred_res = 0; break;
        case 'q': QUIET = 1; break;
        case 'P': pipe = 1; QUIET = 1; break;
        case 1000: prior = 1; break;
        case 1001: fp32 = 1; break;
        case 1002: device = atoi(optarg); break;
        case 1003: seed_opt = atol(optarg); break;
        case 1004: dry_run = 1; break;
        default: return 1;
        }
    }
    argc -= optind;
    argv += optind;
    if (argc == 0 || strcmp(argv[0], "sde")) {
        fprintf(stderr, "%s", help);
        if (argc == 0 || !strcmp(argv[0], "psr") || !strcmp(argv[0], "ode"))
            fprintf(stderr, "error: only the 'sde' implementation is built for the GPU\n");
        return 1;
    }
    if (prior) die("--prior is not supported by this build");

    char str[2048];
    snprintf(str, sizeof str, "%s.done_%d.json", path, general_id);
    unlink(str);

    char err[512];
    pj *settings = pj_load_file("./.settings.json", err, sizeof err);
    if (!settings) die(err);
    char *line = malloc(BUFSIZE);
    if (!fgets(line, BUFSIZE, stdin)) die("no theta on stdin");
    pj *theta = pj_parse(line, err, sizeof err);
    free(line);
    if (!theta) die(err);

    pj *cst = pj_get(settings, "cst"), *orders = pj_get(settings, "orders"), *sdata = pj_get(settings, "data");
    const int N_C = (int)pj_number(pj_get(cst, "N_C"), NULL), N_AC = (int)pj_number(pj_get(cst, "N_AC"), NULL);
    const int N_CAC = N_C * N_AC, N_TS = (int)pj_number(pj_get(cst, "N_TS"), NULL);
    const int N_DATA = (int)pj_number(pj_get(cst, "N_DATA"), NULL);
    const char *model_name = pj_string(pj_get(settings, "model"));
    if (!model_name) die(".settings.json has no 'model' (write it with this build's model builder)");
    int model_id = plom_model_find(model_name);
    if (model_id < 0) die(plom_last_error());
    plom_model_info mi;
    plom_model_info_get(model_id, &mi);
    if (mi.n_cac != N_CAC || mi.n_ts != N_TS) die("settings dimensions differ from the compiled model");

    pj *par_sv = pj_get(orders, "par_sv"), *par_proc = pj_get(orders, "par_proc"), *par_obs = pj_get(orders, "par_obs");
    pj *par_fixed = pj_get(orders, "par_fixed"), *cac_id = pj_get(orders, "cac_id"), *ts_id = pj_get(orders, "ts_id");
    const int n_sv = pj_size(par_sv), n_proc = pj_size(par_proc), n_obs = pj_size(par_obs);
    const int n_par = n_sv + n_proc + n_obs;
    router *R = calloc((size_t)n_par, sizeof(router));
    int ind_theta_remainder = -1;
    for (int i = 0; i < n_par; i++) {
        const pj *lst = i < n_sv ? par_sv : (i < n_sv + n_proc ? par_proc : par_obs);
        int k = i < n_sv ? i : (i < n_sv + n_proc ? i - n_sv : i - n_sv - n_proc);
        const char *key = pj_string(pj_at(lst, k));
        const int is_obs = i >= n_sv + n_proc;
        int missing = pj_get(pj_get(theta, "parameter"), key) == NULL;
        if (missing && i >= n_sv) {
            snprintf(str, sizeof str, "parameter %s missing from theta", key);
            die(str);
        }
        if (missing) ind_theta_remainder = i;
        build_router(&R[i], key, theta, is_obs ? ts_id : cac_id, is_obs ? "time_series_id" : "population_id", missing);
    }
    /* transform_theta then the theta remainder (load.c:400-408, proposal.c:87-114) */
    for (int i = 0; i < n_par; i++)
        for (int g = 0; g < R[i].n_gp; g++) R[i].mean[g] = f_tr(R[i].tr, R[i].guess[g], R[i].min[g], R[i].max[g]);
    for (int cac = 0; cac < N_CAC; cac++) {
        double prop = 0.0;
        for (int i = 0; i < n_sv; i++)
            if (i != ind_theta_remainder) prop += natural(&R[i], R[i].map[cac]);
        if (prop > 1.0) die("sum of the initial conditions without remainder is greater than 1.0");
        if (ind_theta_remainder >= 0) R[ind_theta_remainder].mean[cac] = f_tr(TR_LOGIT, 1.0 - prop, 0.0, 1.0);
    }

    /* covariates: constant per population (value at t = 0), as pop_size_t0 (build.c:1179-1185) */
    pj *pfv = pj_get(sdata, "par_fixed_values");
    const int n_fixed = pj_size(par_fixed);
    double *fixed = calloc((size_t)(n_fixed * N_CAC + 1), sizeof(double));
    for (int k = 0; k < n_fixed; k++) {
        pj *series = pj_get(pfv, pj_string(pj_at(par_fixed, k)));
        for (int cac = 0; cac < N_CAC; cac++) {
            pj *s = pj_at(series, index_of_series(series, pj_string(pj_at(cac_id, cac))));
            pj *y = pj_get(s, "y");
            double y0 = pj_number(pj_at(y, 0), NULL);
            for (int q = 1; q < pj_size(y); q++)
                if (pj_number(pj_at(y, q), NULL) != y0) die("time-varying covariates are not supported by this build yet");
            fixed[k * N_CAC + cac] = y0;
        }
    }
    /* parameter vector in the generated layout: proc [n_proc][N_CAC], obs [n_obs][N_TS],
     * fixed [n_fixed][N_CAC], fixed_obs [..][N_TS], coupling [N_CAC][N_CAC] */
    double *pv = calloc((size_t)mi.pv_size, sizeof(double));
    int o = 0;
    for (int i = n_sv; i < n_sv + n_proc; i++)
        for (int cac = 0; cac < N_CAC; cac++) pv[o++] = natural(&R[i], R[i].map[cac]);
    for (int i = n_sv + n_proc; i < n_par; i++)
        for (int ts = 0; ts < N_TS; ts++) pv[o++] = natural(&R[i], R[i].map[ts]);
    for (int k = 0; k < n_fixed * N_CAC; k++) pv[o++] = fixed[k];
    pj *pfo = pj_get(settings, "par_fixed_obs");
    for (int q = 0; q < pj_size(pfo); q++) {
        int k = index_of(par_fixed, pj_string(pj_at(pfo, q)));
        pj *series = pj_get(pfv, pj_string(pj_at(pfo, q)));
        for (int ts = 0; ts < N_TS; ts++) {
            pj *s = pj_at(series, index_of_series(series, pj_string(pj_at(ts_id, ts))));
            pv[o++] = s ? pj_number(pj_at(pj_get(s, "y"), 0), NULL) : fixed[k * N_CAC];
        }
    }
    pj *coupling = pj_get(pj_get(settings, "coupling"), "matrix");
    for (int a = 0; a < pj_size(coupling); a++)
        for (int b = 0; b < pj_size(pj_at(coupling, a)); b++) pv[o++] = pj_number(pj_at(pj_at(coupling, a), b), NULL);
    if (o != mi.pv_size) {
        snprintf(str, sizeof str, "parameter vector has %d values, the model expects %d", o, mi.pv_size);
        die(str);
    }
    /* initial state: proportions x population size (prediction_util.c:82-115) */
    double *x0 = calloc((size_t)mi.size_proj, sizeof(double));
    for (int i = 0; i < n_sv; i++)
        for (int cac = 0; cac < N_CAC; cac++) x0[i * N_CAC + cac] = natural(&R[i], R[i].map[cac]) * fixed[cac];

    /* data and times (build.c:662-881) */
    int N = (nb_obs > 0 && nb_obs < N_DATA) ? nb_obs : N_DATA;
    double *times = calloc((size_t)N + 1, sizeof(double));
    double *data = calloc((size_t)N * N_TS + 1, sizeof(double));
    pj *jt = pj_get(sdata, "times"), *jd = pj_get(sdata, "data");
    for (int n = 0; n <= N; n++) times[n] = pj_number(pj_at(jt, n), NULL);
    for (int n = 0; n < N; n++)
        for (int ts = 0; ts < N_TS; ts++) {
            pj *v = pj_at(pj_at(jd, n), ts);
            data[n * N_TS + ts] = pj_is_number(v) ? v->num : NAN;
        }
    if (dt <= 0.0) dt = 0.25 / 365.0 * 365.0;
    dt = 1.0 / ((double)round(1.0 / dt));

    if (dry_run) {
        /* the filter inputs as JSON, no GPU call (for tests of the host-side parameter logic) */
        printf("{\"model\": \"%s\", \"N\": %d, \"dt\": %.17g, \"pv\": [", model_name, N, dt);
        for (int k = 0; k < mi.pv_size; k++) printf("%s%.17g", k ? ", " : "", pv[k]);
        printf("], \"x0\": [");
        for (int k = 0; k < mi.size_proj; k++) printf("%s%.17g", k ? ", " : "", x0[k]);
        printf("]}\n");
        return 0;
    }
    if (print_X) fprintf(stderr, "note: the X output (-r) is not written by this build\n");
    const int want_hat = print_hat || (print_pred_res && filter);

    snprintf(str, sizeof str, "Starting plom-smc with the following options: i = %d, J = %d, LIKE_MIN = %g, n_threads = %d",
             general_id, J, like_min, n_threads);
    print_log(str);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);

    plom_ctx *ctx = NULL;
    if (plom_ctx_create(device, &ctx) != PLOM_OK) die(plom_last_error());
    plom_filter_desc d;
    memset(&d, 0, sizeof d);
    d.model_id = model_id;
    d.J = J;
    d.N = N;
    d.precision = fp32 ? PLOM_FP32 : PLOM_FP64;
    d.noises_off = noises_off;
    d.flags = (filter ? 0u : PLOM_FLAG_NO_FILTER) | (want_hat ? PLOM_FLAG_HAT : 0u);
    d.seed = (uint32_t)(seed_opt >= 0 ? seed_opt : 2 + general_id);
    d.stream = 0;
    d.dt = dt;
    d.like_min = like_min;
    plom_filter_inputs in = {times, data, pv, x0};
    plom_filter_outputs out;
    memset(&out, 0, sizeof out);
    double *ess_n = calloc((size_t)N + 1, sizeof(double));
    double *hat = want_hat ? calloc((size_t)(N + 1) * mi.n_hat * 3, sizeof(double)) : NULL;
    double *pres = want_hat ? calloc((size_t)N * N_TS * 2 + 1, sizeof(double)) : NULL;
    out.ess_n = ess_n;
    out.hat = hat;
    out.pred_res = pres;
    if (plom_smc_run(ctx, &d, &in, &out) != PLOM_OK) die(plom_last_error());
    plom_ctx_destroy(ctx);

    clock_gettime(CLOCK_MONOTONIC, &t1);
    double secs = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    if (filter) {
        snprintf(str, sizeof str, "logV: %g, J: %d,  n_all_fail: %d", out.loglik, J, out.n_all_fail);
        print_log(str);
        snprintf(str, sizeof str, "logV_17g: %.17g", out.loglik);
        print_log(str);
    }
    int dd = (int)(secs / 86400.0), hh = (int)((secs - dd * 86400.0) / 3600.0);
    int mm = (int)((secs - dd * 86400.0 - hh * 3600.0) / 60.0);
    snprintf(str, sizeof str, "Done in:= %dd %dh %dm %gs", dd, hh, mm, secs - dd * 86400.0 - hh * 3600.0 - mm * 60.0);
    print_log(str);

    /* trace_<id>.csv (print.c:189-207, 396-440): user-scale values f_inv(mean), log_like */
    if (print_best) {
        snprintf(str, sizeof str, "%strace_%d.csv", path, general_id);
        FILE *ft = fopen(str, "w");
        if (!ft) die("cannot write the trace file");
        fprintf(ft, "index,");
        for (int i = 0; i < n_par; i++)
            if (i != ind_theta_remainder)
                for (int g = 0; g < R[i].n_gp; g++) fprintf(ft, "%s:%s,", R[i].name, R[i].group_name[g]);
        fprintf(ft, "log_like\n0,");
        for (int i = 0; i < n_par; i++)
            if (i != ind_theta_remainder)
                for (int g = 0; g < R[i].n_gp; g++) fprintf(ft, "%g,", f_inv(R[i].tr, R[i].mean[g], R[i].min[g], R[i].max[g]));
        fprintf(ft, "%g\n", out.loglik);
        fclose(ft);
    }
    /* hat_<id>.csv (header_hat, print.c:151-186; print_p_hat, print.c:445-573) */
    const char *remainder = pj_string(pj_get(settings, "remainder"));
    const pj *psz = pj_get(settings, "POP_SIZE_EQ_SUM_SV");
    const int has_rem = !(psz && psz->type == PJ_TRUE) && remainder;
    if (print_hat && hat) {
        snprintf(str, sizeof str, "%shat_%d.csv", path, general_id);
        FILE *fh = fopen(str, "w");
        if (!fh) die("cannot write the hat file");
        fprintf(fh, "time,");
        for (int i = 0; i < n_sv; i++)
            for (int cac = 0; cac < N_CAC; cac++) {
                const char *nm = pj_string(pj_at(par_sv, i)), *cn = pj_string(pj_at(cac_id, cac));
                fprintf(fh, "low95:%s:%s,%s:%s,high95:%s:%s,", nm, cn, nm, cn, nm, cn);
            }
        if (has_rem)
            for (int cac = 0; cac < N_CAC; cac++) {
                const char *cn = pj_string(pj_at(cac_id, cac));
                fprintf(fh, "low95:%s:%s,%s:%s,high95:%s:%s,", remainder, cn, remainder, cn, remainder, cn);
            }
        for (int ts = 0; ts < N_TS; ts++) {
            const char *tn = pj_string(pj_at(ts_id, ts));
            fprintf(fh, "low95:%s,%s,high95:%s%s", tn, tn, tn, ts < N_TS - 1 ? "," : "");
        }
        fprintf(fh, "\n");
        for (int n = 0; n <= N; n++) {
            const double *row = hat + (size_t)n * mi.n_hat * 3;
            fprintf(fh, "%g,", times[n]);
            for (int k = 0; k < mi.n_hat; k++)
                fprintf(fh, "%g,%g,%g%s", row[3 * k], row[3 * k + 1], row[3 * k + 2], k < mi.n_hat - 1 ? "," : "");
            fprintf(fh, "\n");
        }
        fclose(fh);
    }
    /* pred_res_<id>.csv (header_prediction_residuals, print.c:132-143; rows only where the
     * interval has data, print.c:632-701 as called from smc.c:455-457) */
    if (print_pred_res && pres && filter) {
        snprintf(str, sizeof str, "%spred_res_%d.csv", path, general_id);
        FILE *fp = fopen(str, "w");
        if (!fp) die("cannot write the pred_res file");
        fprintf(fp, "time,");
        for (int ts = 0; ts < N_TS; ts++) {
            const char *tn = pj_string(pj_at(ts_id, ts));
            fprintf(fp, "mean:%s,res:%s,", tn, tn);
        }
        fprintf(fp, "ess\n");
        for (int n = 0; n < N; n++) {
            int nonan = 0;
            for (int ts = 0; ts < N_TS; ts++) nonan += !isnan(data[n * N_TS + ts]);
            if (!nonan) continue;
            fprintf(fp, "%g,", times[n + 1]);
            for (int ts = 0; ts < N_TS; ts++) fprintf(fp, "%g,%g,", pres[((size_t)n * N_TS + ts) * 2], pres[((size_t)n * N_TS + ts) * 2 + 1]);
            fprintf(fp, "%g\n", ess_n[n]);
        }
        fclose(fp);
    }
    /* plom_print_done (print.c:33-67): guesses replaced by f_inv(mean) */
    for (int i = 0; i < n_par; i++) {
        if (i == ind_theta_remainder || !R[i].par) continue;
        pj *groups = pj_get(R[i].par, "group");
        for (int g = 0; g < R[i].n_gp; g++) {
            pj *grp = pj_get(groups, R[i].group_name[g]);
            pj *guess = pj_get(grp, "guess");
            if (guess && guess->type == PJ_OBJECT) pj_set(guess, "value", pj_new_real(f_inv(R[i].tr, R[i].mean[g], R[i].min[g], R[i].max[g])));
        }
    }
    if (pipe) {
        pj_dump(theta, stdout);
        fflush(stdout);
    } else {
        snprintf(str, sizeof str, "%s.done_%d.json", path, general_id);
        FILE *fd = fopen(str, "w");
        if (fd) {
            pj_dump(theta, fd);
            fclose(fd);
        }
    }
    if (!QUIET) print_log("clean up...");
    pj_free(theta);
    pj_free(settings);
    return 0;
}
