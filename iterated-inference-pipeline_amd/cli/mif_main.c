/*
 * mif_main.c -- the drop-in `mif` command: maximum likelihood by iterated filtering
 * (Ionides et al. 2006), with the reference's options, updates and outputs
 * (model_builder/C/mif/main.c, mif/mif.c:21-247, mif/methods.c), every pass run by the GPU
 * filter with per-particle parameters (PLOM_FLAG_MIF, include/plom_smc.h).
 *
 *   cat theta.json | ./mif sde -J 10000 -M 50 -a 0.95 -b 2
 *
 * Iteration m (FREEZE = a^(m-1)):
 *   the particles' parameters: theta_j = mean + N(0, (b FREEZE)^2 var_kk) for every
 *     estimated component, redrawn until the initial proportions fit, then followers copy
 *     their leaders; x0_j from theta_j (propose_safe_theta_and_load_X0, proposal.c:40-65);
 *   one filter pass on the device: theta resampled with the state and, after every
 *     observation, moved by N(0, (FREEZE sqrt(delta_t))^2 var_kk) for the estimated
 *     process / observation parameters (initial conditions are resampled only);
 *     per observation the weighted mean of theta and its variance
 *     (mean_var_theta_theoretical_mif, methods.c:146-205), to which delta_t FREEZE^2 var_kk
 *     is added here;
 *   fixed-lag smoothing of the initial conditions at observation L (methods.c:370-388);
 *   the update of the process / observation parameters: the mean over observations of the
 *     filtered means for m <= SWITCH (update_theta_best_stable_mif, methods.c:395-418), the
 *     Ionides (2006) formula after (update_theta_best_king_mif, methods.c:426-461);
 *   a trace row (m, f_inv(mean), log-likelihood of the pass).
 * var is rescaled once (rescale_covariance_mif, methods.c:26-70): process / observation
 * entries by 1/nb_obs, their cross terms with the initial conditions by 1/sqrt(nb_obs).
 *
 * K replicates (--chains K) run in lockstep as one batch (ids id..id+K-1, their own keys
 * and files): one batched pass per iteration for all of them.
 *
 * Random numbers: pass m of replicate k uses the filter key (seed_k, m), seed_k = 2 + id + k;
 * the starting parameters come from Philox-4x32-10 keyed (seed_k, 0xfffffffe), counter
 * (j, m, attempt, block), four Box-Muller normals per block.
 *
 * Implementations 'sde', 'psr' (the default, FP64, initial states rounded) and 'ode' (FP64,
 * adaptive RKF45 under --eps_abs / --eps_rel, every noise off).
 * A "covariance" in theta makes the starting proposals mean + (b FREEZE) Lv z and the random
 * walk mut_fac L z (Lv: Cholesky factor of the rescaled covariance of the estimated
 * components; L: its rows of the process / observation components in component space).
 * --prior (patch_likelihood_prior, methods.c:469-523; is_bayesian, mif/build.c:45): uniform and
 * pseudo-uniform priors move on the logit_ab scale, and at every interval with data the
 * weights take the factor prior^(1/nb_obs) per estimated process / observation component and
 * prior^(1/L) per estimated initial condition while n < L (plom_mif_map.prior*; in the log
 * domain on the device).
 * Differences: the fixed-lag update uses the filtered
 * weighted mean of the initial conditions at the last observation up to L (the reference
 * weights the resampled particles with the weights before resampling, methods.c:370-388,
 * and also re-averages the components that are not estimated); -N is ignored.
 * Extensions: --chains <K>, --lanes <k>, --fp32, --device <k>, --seed <k>, --trace_17g,
 * --dry_run (the pass inputs as JSON, no GPU).
 */
#include <getopt.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "plom_cli.h"
#include "plom_mif.h"
#include "plom_rng.h"

#define START_STREAM 0xfffffffeu

typedef struct {
    int id;
    uint32_t seed;
    double *mean;            /* [n_theta] transformed scale */
    double *theta;           /* [n_theta][J] starting parameters of the pass */
    double *x0;              /* [S][J] */
    double *bart, *vt;       /* [(N+1)][n_theta] */
    double *mm, *mv;         /* [(N+1)][n_theta] device moments */
    double *ess;             /* [N] */
    plom_filter *f;
    FILE *ftrace, *fmif;
    double loglik;
    int naf;
} chain_t;

static double elapsed(const struct timespec *t0)
{
    struct timespec t1;
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0->tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0->tv_nsec);
}

static void fmt_duration(double secs, char *buf, size_t len)
{
    int dd = (int)(secs / 86400.0), hh = (int)((secs - dd * 86400.0) / 3600.0);
    int mm = (int)((secs - dd * 86400.0 - hh * 3600.0) / 60.0);
    snprintf(buf, len, "%dd %dh %dm %gs", dd, hh, mm, secs - dd * 86400.0 - hh * 3600.0 - mm * 60.0);
}

int main(int argc, char *argv[])
{
    int J = 100, general_id = 0, n_threads = 1, fp32 = 0, device = 0, n_chains = 1, lanes = 0;
    int M_iter = 1, SWITCH = 5, ic_only = 0, print_traj = 0, pipe = 0, prior = 0, dry_run = 0;
    long seed_opt = -1;
    double dt = 0.0, like_min = 1e-17, mif_a = 0.975, mif_b = 2.0, prop_L = 0.75;
    unsigned int noises_off = 0;
    double eps_abs = 1e-6, eps_rel = 1e-3;   /* PLOM_EPS_ABS / PLOM_EPS_REL (C/core/plom.h:90-91) */
    char path[1024] = "./";
    static struct option long_options[] = {
        {"prior", no_argument, 0, 1000}, {"no_dem_sto", no_argument, 0, 'x'},
        {"no_white_noise", no_argument, 0, 'y'}, {"no_diff", no_argument, 0, 'z'},
        {"DT", required_argument, 0, 's'}, {"eps_abs", required_argument, 0, 'v'},
        {"eps_rel", required_argument, 0, 'w'}, {"freeze_forcing", required_argument, 0, 'g'},
        {"help", no_argument, 0, 'h'}, {"path", required_argument, 0, 'p'},
        {"id", required_argument, 0, 'i'}, {"n_thread", required_argument, 0, 'N'},
        {"LIKE_MIN", required_argument, 0, 'l'}, {"iter", required_argument, 0, 'M'},
        {"cooling", required_argument, 0, 'a'}, {"heat", required_argument, 0, 'b'},
        {"lag", required_argument, 0, 'L'}, {"switch", required_argument, 0, 'S'},
        {"ic_only", no_argument, 0, 'f'}, {"traj", no_argument, 0, 'r'},
        {"quiet", no_argument, 0, 'q'}, {"pipe", no_argument, 0, 'P'},
        {"chains", required_argument, 0, 1001}, {"lanes", required_argument, 0, 1002},
        {"fp32", no_argument, 0, 1003}, {"device", required_argument, 0, 1004},
        {"seed", required_argument, 0, 1005}, {"trace_17g", no_argument, 0, 1006}, {"dry_run", no_argument, 0, 1007},
        {0, 0, 0, 0}};
    const char *help =
        "PLOM MIF (MI355X)\n"
        "usage:\n"
        "mif [implementation] [--no_dem_sto] [--no_white_noise] [--no_diff]\n"
        "                     [-s, --DT <float>] [-r, --traj] [-p, --path <path>] [-i, --id <integer>]\n"
        "                     [-l, --LIKE_MIN <float>] [-J <integer>] [-M, --iter <integer>]\n"
        "                     [-a, --cooling <float>] [-b, --heat <float>] [-L, --lag <float>] [-S, --switch <integer>]\n"
        "                     [-f --ic_only] [-q, --quiet] [-P, --pipe] [--prior]\n"
        "                     [--chains <K>] [--lanes <k>] [--fp32] [--device <k>] [--seed <k>] [--trace_17g] [-h, --help]\n"
        "where implementation is 'ode', 'sde' or 'psr' (default)\n";
    for (;;) {
        int idx = 0;
        int ch = getopt_long(argc, argv, "rqPhxyzs:v:w:i:J:l:M:a:b:L:S:fp:N:g:", long_options, &idx);
        if (ch == -1) break;
        switch (ch) {
        case 'x': noises_off |= PLOM_NO_DEM_STO_BIT; break;
        case 'y': noises_off |= PLOM_NO_ENV_STO_BIT; break;
        case 'z': noises_off |= PLOM_NO_DRIFT_BIT; break;
        case 's': dt = atof(optarg); break;
        case 'v': eps_abs = atof(optarg); break;
        case 'w': eps_rel = atof(optarg); break;
        case 'g': CLI_FREEZE = atof(optarg); break;
        case 'h': printf("%s", help); return 1;
        case 'p': snprintf(path, sizeof path, "%s", optarg); break;
        case 'N': n_threads = atoi(optarg); break;
        case 'i': general_id = atoi(optarg); break;
        case 'J': J = atoi(optarg); break;
        case 'l': like_min = atof(optarg); break;
        case 'M': M_iter = atoi(optarg); break;
        case 'a': mif_a = atof(optarg); break;
        case 'b': mif_b = atof(optarg); break;
        case 'L': prop_L = atof(optarg); break;
        case 'S': SWITCH = atoi(optarg); break;
        case 'f': ic_only = 1; break;
        case 'r': print_traj = 1; break;
        case 'q': CLI_QUIET = 1; break;
        case 'P': pipe = 1; CLI_QUIET = 1; break;
        case 1000: prior = 1; break;
        case 1001: n_chains = atoi(optarg); break;
        case 1002: lanes = atoi(optarg); break;
        case 1003: fp32 = 1; break;
        case 1004: device = atoi(optarg); break;
        case 1005: seed_opt = atol(optarg); break;
        case 1006: CLI_TRACE_17G = 1; break;
        case 1007: dry_run = 1; break;
        default: return 1;
        }
    }
    argc -= optind;
    argv += optind;
    const int impl = cli_implementation(argc, argv, help);
    uint32_t impl_flags = 0;
    cli_apply_implementation(impl, fp32, &noises_off, &impl_flags);
    const int psr = impl == CLI_PSR;
    if (n_chains < 1 || n_chains > 4096) cli_die("--chains must be in [1, 4096]");
    if (pipe && n_chains > 1) cli_die("-P (pipe) prints one theta: use it with one chain");
    if (J < 2) cli_die("-J must be at least 2 (the variance of theta over the particles)");
    if (M_iter < 0 || M_iter >= (1 << 30)) cli_die("-M must be in [0, 2^30)");

    char str[4096];
    for (int k = 0; k < n_chains; k++) {
        snprintf(str, sizeof str, "%s.done_%d.json", path, general_id + k);
        unlink(str);
    }
    cli_problem P;
    cli_load(&P, prior, -1, dt);
    P.psr = psr;
    P.drift_on = P.n_drift > 0 && !(noises_off & PLOM_NO_DRIFT_BIT) && !(impl_flags & PLOM_FLAG_ODE);
    /* drifting parameters: fitted by fixed-lag smoothing like the initial conditions
     * (p_it_par_sv_and_drift, build.c:745-746); the particles' drift states start at their
     * theta (theta_driftIC2Xdrift, prediction_util.c:119-131) and move with the particle's own
     * volatility */
    int *fls_par = calloc((size_t)P.n_par + 1, sizeof(int));
    for (int i = 0; i < P.n_sv; i++) fls_par[i] = 1;
    for (int i = 0; P.drift_on && i < P.n_drift; i++) fls_par[P.drift_par[i]] = 1;
    const int L = (int)floor(prop_L * P.N_DATA);
    if (ic_only) P.N = (L + 1 > P.N_DATA) ? P.N_DATA : L + 1;
    if (L > P.N) {
        snprintf(str, sizeof str, "L > N_DATA (%d > %d). Please choose a L <= %d", L, P.N, P.N);
        cli_die(str);
    }
    const int n = P.n_theta, N = P.N, S = P.mi.size_proj;
    if (P.mi.n_pp != P.n_proc * P.N_CAC + P.n_obs * P.N_TS) cli_die("the compiled model's per-particle slots differ from the settings");

    /* build_best / load_best: theta, diag(sd_transf^2), estimated components; then
     * rescale_covariance_mif */
    double *theta0 = calloc((size_t)n + 1, sizeof(double));
    cli_theta_from_guess(&P, theta0);
    double *var = calloc((size_t)n * n + 1, sizeof(double));
    int *est = calloc((size_t)n + 1, sizeof(int)), *is_sv = calloc((size_t)n + 1, sizeof(int));
    /* a "covariance" in theta makes the proposals and the random walk multivariate normal
     * (mif.c:84-121, methods.c:275-330) */
    const int is_mvn = cli_load_covariance(&P, var);
    for (int i = 0; i < P.n_par; i++)
        for (int g = 0; g < P.R[i].n_gp; g++) {
            const int o = P.offset[i] + g;
            const int free_par = P.R[i].follow < 0 && !P.R[i].is_remainder;
            if (!is_mvn) {
                const double sdv = free_par ? P.R[i].sd[g] : 0.0;
                var[o * n + o] = sdv > 0.0 ? sdv * sdv : 0.0;
            }
            est[o] = free_par && var[o * n + o] > 0.0;
            is_sv[o] = fls_par[i];
        }
    if (ic_only)
        for (int o = 0; o < n; o++)
            if (!is_sv[o]) est[o] = 0;
    if (noises_off & PLOM_NO_ENV_STO_BIT)
        for (int i = 0; i < P.n_par; i++)
            if (P.noise_par[i])
                for (int g = 0; g < P.R[i].n_gp; g++) est[P.offset[i] + g] = 0;
    /* components not estimated take no part in the covariance */
    for (int o = 0; o < n; o++)
        if (!est[o])
            for (int k = 0; k < n; k++) var[o * n + k] = var[k * n + o] = 0.0;
    for (int r = 0; r < n; r++)
        for (int c2 = 0; c2 < n; c2++) {
            if (!is_sv[r] && !is_sv[c2]) var[r * n + c2] *= 1.0 / (double)N;
            else if (is_sv[r] != is_sv[c2]) var[r * n + c2] *= 1.0 / sqrt((double)N);
        }

    /* the device map: per-particle slots of pv, transforms, jump sizes */
    const int npp = P.mi.n_pp;
    int32_t *pv_src = calloc((size_t)npp + 1, sizeof(int32_t));
    int32_t *mtr = calloc((size_t)n + 1, sizeof(int32_t)), *msc = calloc((size_t)n + 1, sizeof(int32_t));
    int32_t *mdur = calloc((size_t)n + 1, sizeof(int32_t));
    double *mmin = calloc((size_t)n + 1, sizeof(double)), *mmax = calloc((size_t)n + 1, sizeof(double));
    double *mmul = calloc((size_t)n + 1, sizeof(double)), *msd = calloc((size_t)n + 1, sizeof(double));
    int slot = 0;
    for (int i = P.n_sv; i < P.n_sv + P.n_proc; i++)
        for (int cac = 0; cac < P.N_CAC; cac++) pv_src[slot++] = P.offset[i] + P.R[i].map[cac];
    for (int i = P.n_sv + P.n_proc; i < P.n_par; i++)
        for (int ts = 0; ts < P.N_TS; ts++) pv_src[slot++] = P.offset[i] + P.R[i].map[ts];
    for (int i = 0; i < P.n_par; i++)
        for (int g = 0; g < P.R[i].n_gp; g++) {
            const int o = P.offset[i] + g;
            mtr[o] = (int32_t)P.R[i].tr[g];
            msc[o] = (int32_t)P.R[i].sc;
            mdur[o] = P.R[i].is_duration;
            mmin[o] = P.R[i].min[g];
            mmax[o] = P.R[i].max[g];
            mmul[o] = P.R[i].multiplier;
            msd[o] = (est[o] && !is_sv[o]) ? sqrt(var[o * n + o]) : 0.0;
        }
    /* multivariate walk: the Cholesky factor Lv of var restricted to the estimated components
     * (ke), for the proposals; the walk's factor in component space keeps the rows of the
     * process / observation components (the initial conditions are only resampled,
     * methods.c:320-328) */
    int ne = 0;
    int *ke = calloc((size_t)n + 1, sizeof(int));
    for (int o = 0; o < n; o++)
        if (est[o]) ke[ne++] = o;
    double *Lv = NULL, *mchol = NULL;
    if (is_mvn) {
        if (n > PLOM_MIF_MVN_MAX) cli_die("a covariance in theta takes at most 64 components here");
        double *sub = calloc((size_t)ne * ne + 1, sizeof(double));
        Lv = calloc((size_t)ne * ne + 1, sizeof(double));
        for (int a = 0; a < ne; a++)
            for (int b = 0; b < ne; b++) sub[a * ne + b] = var[ke[a] * n + ke[b]];
        if (ne && cli_cholesky(sub, ne, 1.0, Lv)) cli_die("Covariance matrix is not positive definite");
        free(sub);
        mchol = calloc((size_t)n * n + 1, sizeof(double));
        for (int a = 0; a < ne; a++)
            if (!is_sv[ke[a]])
                for (int b = 0; b <= a; b++) mchol[ke[a] * n + ke[b]] = Lv[a * ne + b];
    }
    /* --prior: per estimated component its prior, the exponent and the intervals it applies
     * to (patch_likelihood_prior, methods.c:469-523): 1/nb_obs at every interval for the
     * process / observation parameters, 1/L at intervals n < L for the initial conditions */
    int32_t *mprior = NULL, *muntil = NULL;
    double *mppow = NULL;
    if (prior) {
        mprior = calloc((size_t)n + 1, sizeof(int32_t));
        muntil = calloc((size_t)n + 1, sizeof(int32_t));
        mppow = calloc((size_t)n + 1, sizeof(double));
        for (int i = 0; i < P.n_par; i++)
            for (int g = 0; g < P.R[i].n_gp; g++) {
                const int o = P.offset[i] + g;
                mprior[o] = PLOM_PRIOR_NONE;
                if (!est[o]) continue;
                const cli_prior k = P.R[i].prior[g];
                mprior[o] = k == PRIOR_NORMAL ? PLOM_PRIOR_NORMAL
                            : (k == PRIOR_PSEUDO_UNIFORM ? PLOM_PRIOR_PSEUDO_UNIFORM : PLOM_PRIOR_UNIFORM);
                if (is_sv[o]) {
                    muntil[o] = L;
                    mppow[o] = L > 0 ? 1.0 / (double)L : 0.0;
                } else {
                    muntil[o] = N;
                    mppow[o] = 1.0 / (double)N;
                }
            }
    }
    plom_mif_map mmap = {pv_src, mtr, msc, mdur, mmin, mmax, mmul, msd, mchol, mprior, mppow, muntil};

    /* observations with data, delta_t since the previous one (mif.c:140-150) */
    int *data_rows = calloc((size_t)N + 1, sizeof(int));
    double *delta = calloc((size_t)N + 1, sizeof(double));
    int n_data = 0;
    {
        double acc = 0.0;
        for (int r = 0; r < N; r++) {
            acc += P.times[r + 1] - P.times[r];
            int nonan = 0;
            for (int ts = 0; ts < P.N_TS; ts++) nonan += !isnan(P.data[(size_t)r * P.N_TS + ts]);
            delta[r] = acc;
            if (nonan) { data_rows[n_data++] = r; acc = 0.0; }
        }
    }
    if (n_data == 0) cli_die("no observation with data");
    /* the observation whose filtered initial conditions the fixed-lag smoothing keeps (the
     * reference updates at interval L, mif.c:195-197, which exists when L < N) */
    int last_data_le_L = -1;
    for (int q = 0; q < n_data && L < N; q++)
        if (data_rows[q] <= L) last_data_le_L = data_rows[q];
    if (dry_run) {
        /* the pass inputs as JSON, no GPU call (for tests of the host-side logic) */
        double *pvd = calloc((size_t)P.mi.pv_size, sizeof(double)), *x0d = calloc((size_t)S, sizeof(double));
        cli_pack(&P, theta0, pvd, x0d);
        printf("{\"model\": \"%s\", \"N\": %d, \"L\": %d, \"lag_row\": %d, \"dt\": %.17g, \"theta\": [", P.model_name, N, L,
               last_data_le_L, P.dt);
        for (int o = 0; o < n; o++) printf("%s%.17g", o ? ", " : "", theta0[o]);
        printf("], \"var\": [");
        for (int o = 0; o < n * n; o++) printf("%s%.17g", o ? ", " : "", var[o]);
        printf("], \"estimated\": [");
        for (int o = 0; o < n; o++) printf("%s%d", o ? ", " : "", est[o]);
        printf("], \"sv\": [");
        for (int o = 0; o < n; o++) printf("%s%d", o ? ", " : "", is_sv[o]);
        printf("], \"pv_src\": [");
        for (int k = 0; k < npp; k++) printf("%s%d", k ? ", " : "", pv_src[k]);
        printf("], \"tr\": [");
        for (int o = 0; o < n; o++) printf("%s%d", o ? ", " : "", mtr[o]);
        printf("], \"scale\": [");
        for (int o = 0; o < n; o++) printf("%s%d", o ? ", " : "", msc[o]);
        printf("], \"duration\": [");
        for (int o = 0; o < n; o++) printf("%s%d", o ? ", " : "", mdur[o]);
        printf("], \"min\": [");
        for (int o = 0; o < n; o++) printf("%s%.17g", o ? ", " : "", mmin[o]);
        printf("], \"max\": [");
        for (int o = 0; o < n; o++) printf("%s%.17g", o ? ", " : "", mmax[o]);
        printf("], \"multiplier\": [");
        for (int o = 0; o < n; o++) printf("%s%.17g", o ? ", " : "", mmul[o]);
        printf("], \"sd\": [");
        for (int o = 0; o < n; o++) printf("%s%.17g", o ? ", " : "", msd[o]);
        printf("], \"chol\": [");
        for (int o = 0; mchol && o < n * n; o++) printf("%s%.17g", o ? ", " : "", mchol[o]);
        printf("], \"chol_est\": [");
        for (int o = 0; Lv && o < ne * ne; o++) printf("%s%.17g", o ? ", " : "", Lv[o]);
        printf("], \"prior\": [");
        for (int o = 0; mprior && o < n; o++) printf("%s%d", o ? ", " : "", mprior[o]);
        printf("], \"prior_pow\": [");
        for (int o = 0; mppow && o < n; o++) printf("%s%.17g", o ? ", " : "", mppow[o]);
        printf("], \"prior_until\": [");
        for (int o = 0; muntil && o < n; o++) printf("%s%d", o ? ", " : "", muntil[o]);
        printf("], \"pop0\": [");
        for (int cac = 0; cac < P.N_CAC; cac++) printf("%s%.17g", cac ? ", " : "", P.fixed[cac]);
        printf("], \"drift\": ");
        if (P.drift_on) {
            const int ND = P.n_drift * P.N_CAC;
            const int32_t *ia[5] = {P.dmap.group, P.dmap.tr, P.dmap.scale, P.dmap.duration, P.dmap.vol_slot};
            const char *in[5] = {"group", "tr", "scale", "duration", "vol_slot"};
            const double *da[3] = {P.dmap.min, P.dmap.max, P.dmap.multiplier};
            const char *dn[3] = {"min", "max", "multiplier"};
            printf("{");
            for (int a = 0; a < 5; a++) {
                printf("%s\"%s\": [", a ? ", " : "", in[a]);
                for (int q = 0; q < ND; q++) printf("%s%d", q ? ", " : "", ia[a][q]);
                printf("]");
            }
            for (int a = 0; a < 3; a++) {
                printf(", \"%s\": [", dn[a]);
                for (int q = 0; q < ND; q++) printf("%s%.17g", q ? ", " : "", da[a][q]);
                printf("]");
            }
            printf("}");
        } else {
            printf("null");
        }
        printf(", \"delta\": [");
        for (int r = 0; r < N; r++) printf("%s%.17g", r ? ", " : "", delta[r]);
        printf("], \"data_rows\": [");
        for (int q = 0; q < n_data; q++) printf("%s%d", q ? ", " : "", data_rows[q]);
        printf("], \"times0\": %.17g, \"pv\": [", P.times[0]);
        for (int k = 0; k < P.mi.pv_size; k++) printf("%s%.17g", k ? ", " : "", pvd[k]);
        printf("], \"x0\": [");
        for (int k = 0; k < S; k++) printf("%s%.17g", k ? ", " : "", x0d[k]);
        printf("]}\n");
        return 0;
    }

    snprintf(str, sizeof str, "Starting plom-mif with the following options: i = %d, J = %d, LIKE_MIN = %g, M = %d, a = %g, b = %g, L = %g, SWITCH = %d, N_THREADS = %d",
             general_id, J, like_min, M_iter, mif_a, mif_b, prop_L, SWITCH, n_threads);
    cli_log(str);
    struct timespec t_all;
    clock_gettime(CLOCK_MONOTONIC, &t_all);

    plom_ctx *ctx = NULL;
    if (plom_ctx_create(device, &ctx) != PLOM_OK) cli_die(plom_last_error());
    const uint32_t seed0 = (uint32_t)(seed_opt >= 0 ? seed_opt : 2 + general_id);
    chain_t *C = calloc((size_t)n_chains, sizeof(chain_t));
    plom_filter **fs = calloc((size_t)n_chains, sizeof(plom_filter *));
    double *pv = calloc((size_t)P.mi.pv_size, sizeof(double)), *x0one = calloc((size_t)S, sizeof(double));
    cli_pack(&P, theta0, pv, x0one);
    for (int k = 0; k < n_chains; k++) {
        chain_t *c = &C[k];
        c->id = general_id + k;
        c->seed = seed0 + (uint32_t)k;
        c->mean = calloc((size_t)n + 1, sizeof(double));
        memcpy(c->mean, theta0, (size_t)n * sizeof(double));
        c->theta = calloc((size_t)n * J, sizeof(double));
        c->x0 = calloc((size_t)S * J, sizeof(double));
        c->bart = calloc((size_t)(N + 1) * n, sizeof(double));
        c->vt = calloc((size_t)(N + 1) * n, sizeof(double));
        c->mm = calloc((size_t)(N + 1) * n, sizeof(double));
        c->mv = calloc((size_t)(N + 1) * n, sizeof(double));
        c->ess = calloc((size_t)N + 1, sizeof(double));
        c->ftrace = cli_fopen(path, "trace", c->id);
        cli_trace_header(&P, c->ftrace);
        cli_trace_row(&P, c->ftrace, 0, c->mean, NAN);
        fflush(c->ftrace);
        if (print_traj) {
            c->fmif = cli_fopen(path, "mif", c->id);
            fprintf(c->fmif, "index,time,");
            for (int i = 0; i < P.n_par; i++)
                if (i != P.ind_theta_remainder)
                    for (int g = 0; g < P.R[i].n_gp; g++)
                        fprintf(c->fmif, "mean:%s:%s,var:%s:%s,", P.R[i].name, P.R[i].group_name[g], P.R[i].name, P.R[i].group_name[g]);
            fprintf(c->fmif, "ess\n");
        }
        plom_filter_desc d;
        memset(&d, 0, sizeof d);
        d.model_id = P.model_id;
        d.J = J;
        d.N = N;
        d.precision = fp32 ? PLOM_FP32 : PLOM_FP64;
        d.noises_off = noises_off;
        d.eps_abs = eps_abs;
        d.eps_rel = eps_rel;
        d.flags = PLOM_FLAG_MIF | impl_flags;
        d.seed = c->seed;
        d.dt = P.dt;
        d.like_min = like_min;
        d.n_theta = n;
        plom_filter_inputs in = {P.times, P.data, pv, x0one, &mmap, P.has_cov ? P.cov : NULL,
                                  P.drift_on ? &P.dmap : NULL};
        if (plom_filter_create(ctx, &d, &in, &fs[k]) != PLOM_OK) cli_die(plom_last_error());
        c->f = fs[k];
    }
    plom_batch *batch = NULL;
    if (plom_batch_create(ctx, n_chains, fs, &batch) != PLOM_OK) cli_die(plom_last_error());
    if (lanes <= 0) lanes = n_chains < 4 ? n_chains : 4;
    if (plom_batch_configure(ctx, batch, lanes) != PLOM_OK) cli_die(plom_last_error());
    double *mut = calloc((size_t)N + 1, sizeof(double));
    double *th = calloc((size_t)n + 4, sizeof(double));
    double *zs = calloc((size_t)n + 4, sizeof(double));

    for (int m = 1; m <= M_iter; m++) {
        struct timespec t_it;
        clock_gettime(CLOCK_MONOTONIC, &t_it);
        const double FREEZE = pow(mif_a, (double)(m - 1));
        for (int r = 0; r < N; r++) mut[r] = FREEZE * sqrt(delta[r]);
        const double sd_fac = mif_b * FREEZE;
        for (int k = 0; k < n_chains; k++) {
            chain_t *c = &C[k];
            /* fill_theta_bart_and_Vt_mif (methods.c:114-138) */
            for (int r = 0; r <= N; r++)
                for (int o = 0; o < n; o++) {
                    c->bart[(size_t)r * n + o] = c->mean[o];
                    c->vt[(size_t)r * n + o] = r == 0 ? (sd_fac * sd_fac) * var[o * n + o] : 0.0;
                }
            /* the particles' starting parameters (ran_proposal, proposal.c:71-84) */
            for (int j = 0; j < J; j++) {
                for (int attempt = 0;; attempt++) {
                    if (attempt > 1000000) cli_die("no proposal satisfies the initial-condition constraint");
                    memcpy(th, c->mean, (size_t)n * sizeof(double));
                    /* ne normals, four per Philox draw (ctr = (j, m, attempt, block)) */
                    for (int q = 0; q < ne; q += 4) {
                        uint32_t w[4];
                        const uint32_t ctr[4] = {(uint32_t)j, (uint32_t)m, (uint32_t)attempt, (uint32_t)(q >> 2)};
                        const uint32_t key[2] = {c->seed, START_STREAM};
                        plom_philox4x32_10(ctr, key, w);
                        plom_bm_pair(w[0], w[1], plom_bm_tab, &zs[q], &zs[q + 1]);
                        plom_bm_pair(w[2], w[3], plom_bm_tab, &zs[q + 2], &zs[q + 3]);
                    }
                    for (int q = 0; q < ne; q++) {
                        const int o = ke[q];
                        if (is_mvn) {
                            /* ran_proposal_chol (proposal.c:40-69): mean + (b FREEZE) Lv z */
                            double acc = 0.0;
                            for (int p = 0; p <= q; p++) acc = acc + Lv[q * ne + p] * zs[p];
                            th[o] = c->mean[o] + sd_fac * acc;
                        } else {
                            th[o] = c->mean[o] + sd_fac * sqrt(var[o * n + o]) * zs[q];
                        }
                    }
                    if (cli_theta_remainder(&P, th) == 0) break;
                }
                cli_follow(&P, th);
                for (int o = 0; o < n; o++) c->theta[(size_t)o * J + j] = th[o];
                for (int i = 0; i < P.n_sv; i++)
                    for (int cac = 0; cac < P.N_CAC; cac++)
                    {
                        double v = cli_natural(&P, i, P.R[i].map[cac], th) * P.fixed[cac];
                        if (psr) v = round(v);
                        c->x0[(size_t)(i * P.N_CAC + cac) * J + j] = v;
                    }
                for (int i = 0; P.drift_on && i < P.n_drift; i++)
                    for (int cac = 0; cac < P.N_CAC; cac++) {
                        const int ip = P.drift_par[i];
                        c->x0[(size_t)(P.mi.drift_offset + i * P.N_CAC + cac) * J + j] = th[P.offset[ip] + P.R[ip].map[cac]];
                    }
            }
            if (plom_filter_mif_start(ctx, c->f, c->theta, c->x0, mut, c->seed, (uint32_t)m) != PLOM_OK) cli_die(plom_last_error());
        }
        if (plom_batch_launch(ctx, batch, 1) != PLOM_OK) cli_die(plom_last_error());
        for (int k = 0; k < n_chains; k++) {
            chain_t *c = &C[k];
            plom_filter_outputs out;
            memset(&out, 0, sizeof out);
            out.mif_mean = c->mm;
            out.mif_var = c->mv;
            out.ess_n = c->ess;
            if (plom_filter_read(ctx, c->f, &out) != PLOM_OK) cli_die(plom_last_error());
            c->loglik = out.loglik;
            c->naf = out.n_all_fail;
            /* the filtered means and variances of the observations (methods.c:146-205) */
            for (int q = 0; q < n_data; q++) {
                const int r = data_rows[q];
                const double var_fac = delta[r] * (FREEZE * FREEZE);
                for (int o = 0; o < n; o++) {
                    if (!est[o] || (is_sv[o] && !print_traj)) continue;
                    double v = c->mv[(size_t)(r + 1) * n + o];
                    if (v < 0.0 || isinf(v) || isnan(v)) v = 0.0;
                    c->bart[(size_t)(r + 1) * n + o] = c->mm[(size_t)(r + 1) * n + o];
                    c->vt[(size_t)(r + 1) * n + o] = v + var_fac * var[o * n + o];
                }
                if (c->fmif) {
                    fprintf(c->fmif, "%d,%d,", m, (int)P.times[r + 1]);
                    for (int i = 0; i < P.n_par; i++)
                        if (i != P.ind_theta_remainder)
                            for (int g = 0; g < P.R[i].n_gp; g++) {
                                const int o = P.offset[i] + g;
                                fprintf(c->fmif, "%g,%g,", c->bart[(size_t)(r + 1) * n + o], c->vt[(size_t)(r + 1) * n + o]);
                            }
                    fprintf(c->fmif, "%g\n", c->ess[r]);
                }
            }
            /* fixed-lag smoothing of the estimated initial conditions */
            if (last_data_le_L >= 0)
                for (int o = 0; o < n; o++)
                    if (est[o] && is_sv[o]) c->mean[o] = c->mm[(size_t)(last_data_le_L + 1) * n + o];
            /* update of the process / observation parameters */
            for (int o = 0; o < n; o++) {
                if (!est[o] || is_sv[o]) continue;
                if (m <= SWITCH) {
                    double tmp = 0.0;
                    for (int q = 0; q < n_data; q++) tmp += c->bart[(size_t)(data_rows[q] + 1) * n + o];
                    c->mean[o] = tmp / (double)n_data;
                } else {
                    int nnp1 = data_rows[0];
                    double tmp = (c->bart[(size_t)(nnp1 + 1) * n + o] - c->bart[o]) / c->vt[(size_t)(nnp1 + 1) * n + o];
                    for (int q = 1; q < n_data; q++) {
                        const int nn = data_rows[q - 1];
                        nnp1 = data_rows[q];
                        tmp += (c->bart[(size_t)(nnp1 + 1) * n + o] - c->bart[(size_t)(nn + 1) * n + o]) /
                               c->vt[(size_t)(nnp1 + 1) * n + o];
                    }
                    c->mean[o] = c->mean[o] + ((P.times[0] + mif_b * mif_b) * FREEZE * FREEZE * var[o * n + o] * tmp);
                }
            }
            char dur[128], pre[64] = "";
            if (n_chains > 1) snprintf(pre, sizeof pre, "chain %d: ", c->id);
            fmt_duration(elapsed(&t_it), dur, sizeof dur);
            snprintf(str, sizeof str, "%siteration number:%d\t logV: %g\t n_all_fail: %d\t computed in:= %s", pre, m, c->loglik,
                     c->naf, dur);
            cli_log(str);
            cli_trace_row(&P, c->ftrace, m, c->mean, c->loglik);
            fflush(c->ftrace);
        }
    }
    for (int k = 0; k < n_chains; k++) {
        fclose(C[k].ftrace);
        if (C[k].fmif) fclose(C[k].fmif);
        cli_done(&P, C[k].mean, path, C[k].id, pipe);
    }
    plom_batch_destroy(ctx, batch);
    for (int k = 0; k < n_chains; k++) plom_filter_destroy(ctx, fs[k]);
    plom_ctx_destroy(ctx);
    char dur[128];
    fmt_duration(elapsed(&t_all), dur, sizeof dur);
    snprintf(str, sizeof str, "Done in:= %s", dur);
    cli_log(str);
    cli_log("clean up...");
    return 0;
}
