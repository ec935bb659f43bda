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
 * Parameter handling (routers, transforms, theta remainder, initial state) is the shared
 * host logic of plom_cli.c.
 *
 * Implementations "sde", "psr" (the default, as main.c:223-238: FP64, the initial state
 * rounded to counts, prediction_util.c:108-110) and "ode" (FP64, adaptive RKF45 under
 * --eps_abs / --eps_rel, every noise off).
 * Differences: hat, pred_res and the X output of -r use this build's
 * arithmetic (exact integer weights for the 95 % bounds, tree sums);
 * -N is accepted and ignored (the GPU replaces threads).
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

#include "plom_cli.h"

/* ------------------------------------------------------------------ main */
int main(int argc, char *argv[])
{
    int filter = 1, J = 1, nb_obs = -1, general_id = 0, n_threads = 1, fp32 = 0, device = 0;
    int print_best = 1, print_hat = 1, print_pred_res = 1, print_X = 0, pipe = 0, prior = 0, dry_run = 0;
    long seed_opt = -1;
    double dt = 0.0, like_min = 1e-17;
    unsigned int noises_off = 0;
    double eps_abs = 1e-6, eps_rel = 1e-3;   /* PLOM_EPS_ABS / PLOM_EPS_REL (C/core/plom.h:90-91) */
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
        "where implementation is 'ode', 'sde' or 'psr' (default)\n";
    for (;;) {
        int idx = 0;
        int ch = getopt_long(argc, argv, "qPxyzs:v:w:p:i:J:l:trbehrN:o:g:d", long_options, &idx);
        if (ch == -1) break;
        switch (ch) {
        case 'x': noises_off |= PLOM_NO_DEM_STO_BIT; break;
        case 'y': noises_off |= PLOM_NO_ENV_STO_BIT; break;
        case 'z': noises_off |= PLOM_NO_DRIFT_BIT; break;
        case 's': dt = atof(optarg); break;
        case 'v': eps_abs = atof(optarg); break;
        case 'w': eps_rel = atof(optarg); break;
        case 'g': CLI_FREEZE = atof(optarg); break;
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
        case 'q': CLI_QUIET = 1; break;
        case 'P': pipe = 1; CLI_QUIET = 1; break;
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
    const int impl = cli_implementation(argc, argv, help);
    uint32_t impl_flags = 0;
    cli_apply_implementation(impl, fp32, &noises_off, &impl_flags);
    const int psr = impl == CLI_PSR;

    char str[2048];
    snprintf(str, sizeof str, "%s.done_%d.json", path, general_id);
    unlink(str);

    cli_problem P;
    /* with --prior, parameters with a uniform or pseudo-uniform prior use the logit_ab
     * transform (is_bayesian, smc/main.c:246) */
    cli_load(&P, prior, nb_obs, dt);
    P.psr = psr;
    P.drift_on = P.n_drift > 0 && !(noises_off & PLOM_NO_DRIFT_BIT) && !(impl_flags & PLOM_FLAG_ODE);
    const int N = P.N, N_TS = P.N_TS;
    const plom_model_info mi = P.mi;
    double *theta = calloc((size_t)P.n_theta + 1, sizeof(double));
    cli_theta_from_guess(&P, theta);
    double *pv = calloc((size_t)mi.pv_size, sizeof(double));
    double *x0 = calloc((size_t)mi.size_proj, sizeof(double));
    cli_pack(&P, theta, pv, x0);
    const double *times = P.times, *data = P.data;
    dt = P.dt;
    const char *model_name = P.model_name;

    if (dry_run) {
        /* the filter inputs as JSON, no GPU call (for tests of the host-side parameter logic) */
        printf("{\"model\": \"%s\", \"N\": %d, \"dt\": %.17g, \"pv\": [", model_name, N, dt);
        for (int k = 0; k < mi.pv_size; k++) printf("%s%.17g", k ? ", " : "", pv[k]);
        printf("], \"x0\": [");
        for (int k = 0; k < mi.size_proj; k++) printf("%s%.17g", k ? ", " : "", x0[k]);
        printf("]}\n");
        return 0;
    }
    const int want_hat = print_hat || (print_pred_res && filter);

    snprintf(str, sizeof str, "Starting plom-smc with the following options: i = %d, J = %d, LIKE_MIN = %g, n_threads = %d",
             general_id, J, like_min, n_threads);
    cli_log(str);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);

    plom_ctx *ctx = NULL;
    if (plom_ctx_create(device, &ctx) != PLOM_OK) cli_die(plom_last_error());
    plom_filter_desc d;
    memset(&d, 0, sizeof d);
    d.model_id = P.model_id;
    d.J = J;
    d.N = N;
    d.precision = fp32 ? PLOM_FP32 : PLOM_FP64;
    d.noises_off = noises_off;
    d.eps_abs = eps_abs;
    d.eps_rel = eps_rel;
    d.flags = (filter ? 0u : PLOM_FLAG_NO_FILTER) | (want_hat ? PLOM_FLAG_HAT : 0u) | (print_X ? PLOM_FLAG_ALL_STATES : 0u) |
              impl_flags;
    d.seed = (uint32_t)(seed_opt >= 0 ? seed_opt : 2 + general_id);
    d.stream = 0;
    d.dt = dt;
    d.like_min = like_min;
    plom_filter_inputs in = {times, data, pv, x0, NULL, P.has_cov ? P.cov : NULL,
                              P.drift_on ? &P.dmap : NULL};
    plom_filter_outputs out;
    memset(&out, 0, sizeof out);
    double *ess_n = calloc((size_t)N + 1, sizeof(double));
    double *hat = want_hat ? calloc((size_t)(N + 1) * mi.n_hat * 3, sizeof(double)) : NULL;
    double *pres = want_hat ? calloc((size_t)N * N_TS * 2 + 1, sizeof(double)) : NULL;
    out.ess_n = ess_n;
    out.hat = hat;
    out.pred_res = pres;
    /* -r: every particle after every interval (print_X, smc.c:463-465) */
    double *xall = print_X ? calloc((size_t)(N + 1) * J * mi.size_proj + 1, sizeof(double)) : NULL;
    double *xobs = print_X ? calloc((size_t)(N + 1) * J * N_TS * 2 + 1, sizeof(double)) : NULL;
    if (print_X && (!xall || !xobs)) cli_die("not enough host memory for the X output (-r)");
    out.all_states = xall;
    out.all_obs = xobs;
    if (plom_smc_run(ctx, &d, &in, &out) != PLOM_OK) cli_die(plom_last_error());
    plom_ctx_destroy(ctx);
    /* --prior: the log prior of theta joins the log-likelihood (smc/main.c:335-345) */
    if (prior) {
        double lp;
        cli_log_prior(&P, theta, like_min, &lp);
        out.loglik += lp;
    }

    clock_gettime(CLOCK_MONOTONIC, &t1);
    double secs = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    if (filter) {
        snprintf(str, sizeof str, "logV: %g, J: %d,  n_all_fail: %d", out.loglik, J, out.n_all_fail);
        cli_log(str);
        snprintf(str, sizeof str, "logV_17g: %.17g", out.loglik);
        cli_log(str);
    }
    int dd = (int)(secs / 86400.0), hh = (int)((secs - dd * 86400.0) / 3600.0);
    int mm = (int)((secs - dd * 86400.0 - hh * 3600.0) / 60.0);
    snprintf(str, sizeof str, "Done in:= %dd %dh %dm %gs", dd, hh, mm, secs - dd * 86400.0 - hh * 3600.0 - mm * 60.0);
    cli_log(str);

    /* trace_<id>.csv (print.c:189-207, 396-440): user-scale values f_inv(theta), log_like */
    if (print_best) {
        FILE *ft = cli_fopen(path, "trace", general_id);
        cli_trace_header(&P, ft);
        cli_trace_row(&P, ft, 0, theta, out.loglik);
        fclose(ft);
    }
    /* X_<id>.csv (header_X, print.c:86-129; print_X, print.c:360-392): row 0 then every interval,
     * every particle of the resampled cloud */
    if (print_X) {
        FILE *fx = cli_fopen(path, "X", general_id);
        cli_x_header(&P, fx);
        for (int n = 0; n <= N; n++)
            for (int j = 0; j < J; j++)
                cli_x_row(&P, fx, j, n, xall + ((size_t)n * J + j) * mi.size_proj, xobs + ((size_t)n * J + j) * N_TS * 2);
        fclose(fx);
    }
    /* hat_<id>.csv (header_hat, print.c:151-186; print_p_hat, print.c:445-573) */
    const pj *par_sv = P.par_sv, *cac_id = P.cac_id, *ts_id = P.ts_id;
    const int n_sv = P.n_sv, N_CAC = P.N_CAC;
    if (print_hat && hat) {
        FILE *fh = cli_fopen(path, "hat", general_id);
        fprintf(fh, "time,");
        for (int i = 0; i < n_sv; i++)
            for (int cac = 0; cac < N_CAC; cac++) {
                const char *nm = pj_string(pj_at(par_sv, i)), *cn = pj_string(pj_at(cac_id, cac));
                fprintf(fh, "low95:%s:%s,%s:%s,high95:%s:%s,", nm, cn, nm, cn, nm, cn);
            }
        if (P.has_rem)
            for (int cac = 0; cac < N_CAC; cac++) {
                const char *cn = pj_string(pj_at(cac_id, cac)), *rn = P.remainder;
                fprintf(fh, "low95:%s:%s,%s:%s,high95:%s:%s,", rn, cn, rn, cn, rn, cn);
            }
        for (int ts = 0; ts < N_TS; ts++) {
            const char *tn = pj_string(pj_at(ts_id, ts));
            fprintf(fh, "low95:%s,%s,high95:%s%s", tn, tn, tn, ts < N_TS - 1 ? "," : "");
        }
        /* drift: per group of each drifting parameter (header_hat, print.c:175-183) */
        for (int i = 0; P.drift_on && i < P.n_drift; i++) {
            const cli_router *r = &P.R[P.drift_par[i]];
            for (int g = 0; g < r->n_gp; g++)
                fprintf(fh, ",low95:diff:%s:%s,diff:%s:%s,high95:diff:%s:%s", r->name, r->group_name[g], r->name,
                        r->group_name[g], r->name, r->group_name[g]);
        }
        fprintf(fh, "\n");
        const int nh_base = mi.n_hat - mi.n_drift * N_CAC;
        for (int n = 0; n <= N; n++) {
            const double *row = hat + (size_t)n * mi.n_hat * 3;
            fprintf(fh, "%g,", times[n]);
            for (int k = 0; k < nh_base; k++)
                fprintf(fh, "%g,%g,%g%s", row[3 * k], row[3 * k + 1], row[3 * k + 2], k < nh_base - 1 ? "," : "");
            /* the drift entry of a group is that of its first population */
            for (int i = 0; P.drift_on && i < P.n_drift; i++) {
                const cli_router *r = &P.R[P.drift_par[i]];
                for (int g = 0; g < r->n_gp; g++)
                    for (int c = 0; c < N_CAC; c++)
                        if (r->map[c] == g) {
                            const int k = nh_base + i * N_CAC + c;
                            fprintf(fh, ",%g,%g,%g", row[3 * k], row[3 * k + 1], row[3 * k + 2]);
                            break;
                        }
            }
            fprintf(fh, "\n");
        }
        fclose(fh);
    }
    /* pred_res_<id>.csv (header_prediction_residuals, print.c:132-143; rows only where the
     * interval has data, print.c:632-701 as called from smc.c:455-457) */
    if (print_pred_res && pres && filter) {
        FILE *fp = cli_fopen(path, "pred_res", general_id);
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
    /* plom_print_done (print.c:33-67): guesses replaced by f_inv(theta) */
    cli_done(&P, theta, path, general_id, pipe);
    cli_log("clean up...");
    pj_free(P.theta);
    pj_free(P.settings);
    return 0;
}
