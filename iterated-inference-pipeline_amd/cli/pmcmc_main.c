/*
 * pmcmc_main.c -- the drop-in `pmcmc` command: particle marginal Metropolis-Hastings with
 * the reference's options, adaptation and outputs (model_builder/C/pmcmc/main.c,
 * pmcmc/pmcmc.c:40-266), every likelihood evaluated by the GPU particle filter through
 * the C ABI (include/plom_smc.h).
 *
 *   cat theta.json | ./pmcmc sde -J 1000 -M 10000 --full
 *
 * K independent chains (--chains K, default 1) run in lockstep: chain k is what the
 * reference's `pmcmc -i <id + k>` process would be (its own random key, its own output
 * files), and every iteration evaluates the K proposals as one batch -- the filters are
 * updated in place (plom_filter_update) and the captured launch graph of the batch is
 * replayed, so an iteration costs one batched pass over the data on the device.
 *
 * Algorithm (one chain):
 *   iteration 0: filter at the initial theta, accepted; trace row 0; X rows (index 0).
 *   iteration m >= 1:
 *     increment_iteration_counters (mcmc_util.c:163-182);
 *     proposal scale (get_var_and_sd_fac, mcmc_util.c:88-147): with --full,
 *       epsilon <- epsilon exp(a^(m-1) (ar - 0.234)) while m > E and m ar < S, else 1,
 *       capped at epsilon_max; sd_fac = epsilon 2.38 / sqrt(n); the empirical covariance
 *       replaces the initial one once m ar >= S.  Sequential: sd_fac = 1 and the order of
 *       the n estimated components is shuffled at the start of every cycle;
 *     proposal (propose_safe_theta_and_load_X0, proposal.c:40-65): MVN around the current
 *       theta (--full, mvn.c:105-150) or one component (mcmc_util.c:186-197), redrawn until
 *       the initial proportions sum to at most 1, then followers copy their leaders;
 *     filter; Metropolis-Hastings (metropolis_hastings_prior.c:24-50) with the proposal
 *       density's change-of-variable factor and the priors on the natural scale
 *       (metropolis_hastings_prior.c:60-92, proposal.c:120-188);
 *     acceptance rates (mcmc_util.c:156-166); after every full cycle the empirical
 *       covariance (mvn.c:196-233), the X rows of the last filter (thinned), the trace row
 *       and the --acc row.
 *   At the end: covariance_<id>.csv (mcmc_util.c:283-348) and the theta with the last
 *   accepted values (.done_<id>.json, or stdout with -P).
 *
 * Random numbers: the filter of chain k at iteration m uses the key (seed_k, m) with
 * seed_k = 2 + id + k (build.c:965-971), so iteration 0 of chain 0 is the `smc` filter of
 * the same theta and id.  The chain's own draws come from Philox-4x32-10 keyed
 * (seed_k, 0xffffffff): counter (m, 1, attempt, block) for the proposal normals (Box-Muller,
 * four per block), (m, 2, 0, 0) for the acceptance uniform, (m, 3, block, 0) for the
 * shuffle.  Filters never use stream 0xffffffff (iterations are below 2^31).
 *
 * Implementations 'sde', 'psr' (the default, FP64, initial states rounded) and 'ode' (FP64,
 * adaptive RKF45 under --eps_abs / --eps_rel, every noise off).
 * Differences: -Z/-c (zmq workers) and -N are accepted
 * and ignored; per-group transforms are used throughout (the reference takes group 0's
 * f_inv in the trace, the done file and the proposal's change of variable); the proposal
 * density is evaluated in the log domain (same ratio); a non positive definite
 * empirical covariance falls back to the initial one for that proposal (with a warning)
 * where the reference's GSL error handler would abort.
 * Extensions: --chains <K>, --lanes <k> (streams of the batch), --fp32, --device <k>,
 * --seed <k> (seed_0), --trace_17g (trace values with 17 significant digits), --dry_run
 * (print the initial theta, covariance and filter inputs as JSON and stop before the GPU).
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
#include "plom_rng.h"

#define CHAIN_STREAM 0xffffffffu
#define PURPOSE_PROPOSAL 1u
#define PURPOSE_ACCEPT 2u
#define PURPOSE_SHUFFLE 3u

static double LIKE_MIN = 1e-17;

typedef struct {
    int id;
    uint32_t seed;
    double *mean, *proposed;        /* [n_theta] transformed scale */
    double *mean_sampling;          /* [n_theta] */
    double *var_sampling;           /* [n_theta^2] empirical covariance */
    int *tbe;                       /* order of the estimated components */
    double *ar, *sm_ar;             /* [n_theta] per-component acceptance rates */
    double global_ar, sm_global_ar, epsilon;
    int has_cycled, m_full, cycle_id;
    double L_prev, L_new;
    int accept;
    plom_filter *f;
    FILE *ftrace, *fx, *facc;
    double *pv, *x0;
    struct timespec t_begin;
} chain_t;

typedef struct {
    const cli_problem *P;
    int n;                          /* n_theta */
    double *var0;                   /* [n^2] initial covariance (transformed scale) */
    int *is_estimated;              /* [n] */
    int n_tbe;
    int *tbe0;                      /* estimated components, in offset order */
    cli_tr *tr;                     /* [n] transform of each component */
    double *pmin, *pmax;            /* [n] prior bounds */
    cli_prior *prior;               /* [n] */
    int full;
    double a, epsilon_max, alpha;
    int m_switch, m_eps, smooth;
} mcmc_t;

/* ------------------------------------------------------------------ random numbers */
static void chain_words(const chain_t *c, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t w[4])
{
    const uint32_t ctr[4] = {c0, c1, c2, c3};
    const uint32_t key[2] = {c->seed, CHAIN_STREAM};
    plom_philox4x32_10(ctr, key, w);
}

/* n standard normals for attempt `attempt` of iteration m */
static void chain_normals(const chain_t *c, int m, int attempt, int n, double *z)
{
    for (int b = 0; 4 * b < n; b++) {
        uint32_t w[4];
        double y[4];
        chain_words(c, (uint32_t)m, PURPOSE_PROPOSAL, (uint32_t)attempt, (uint32_t)b, w);
        plom_bm_pair(w[0], w[1], plom_bm_tab, &y[0], &y[1]);
        plom_bm_pair(w[2], w[3], plom_bm_tab, &y[2], &y[3]);
        for (int k = 0; k < 4 && 4 * b + k < n; k++) z[4 * b + k] = y[k];
    }
}

/* ------------------------------------------------------------------ priors and densities */
/* log_prob_prior: returns 0 on success, 1 on a numerical failure (the value is still summed) */
static int log_prob_prior(const mcmc_t *M, const double *theta, double *out)
{
    int err = 0;
    double Lp = 0.0;
    for (int k = 0; k < M->n; k++) {
        if (!M->is_estimated[k]) continue;
        const double x = cli_f_inv(M->tr[k], theta[k], M->pmin[k], M->pmax[k]);
        double p = cli_prior_pdf(M->prior[k], x, M->pmin[k], M->pmax[k]);
        if (isnan(p) || isinf(p) || p < 0.0) {
            err = 1;
            p = LIKE_MIN;
        } else if (p <= LIKE_MIN) {
            p = LIKE_MIN;
        }
        Lp += log(p);
    }
    *out = Lp;
    return err;
}

/* log q(a | b) on the natural scale: the transformed-scale density divided by the
 * derivative of f_inv at a (log_prob_proposal, proposal.c:120-188).  With mvn, Lc is the
 * Cholesky factor of sd_fac^2 var restricted to the estimated components. */
static int log_prob_proposal(const mcmc_t *M, const double *a, const double *b, const double *var, double sd_fac,
                             const double *Lc, double *out)
{
    double Lp = 0.0;
    const int n = M->n_tbe;
    if (M->full) {
        /* log N(a; b, L L^T) = -n/2 log 2 pi - sum log L_ii - |L^-1 (a - b)|^2 / 2 */
        double *y = calloc((size_t)n + 1, sizeof(double));
        double q = 0.0, ld = 0.0;
        for (int i = 0; i < n; i++) {
            double v = a[M->tbe0[i]] - b[M->tbe0[i]];
            for (int k = 0; k < i; k++) v -= Lc[i * n + k] * y[k];
            y[i] = v / Lc[i * n + i];
            q += y[i] * y[i];
            ld += log(Lc[i * n + i]);
        }
        free(y);
        Lp = -0.5 * n * log(2.0 * M_PI) - ld - 0.5 * q;
    }
    for (int k = 0; k < M->n; k++) {
        if (!M->is_estimated[k]) continue;
        const double der = cli_f_inv_der(M->tr[k], a[k], M->pmin[k], M->pmax[k]);
        if (!M->full) {
            const double p = cli_gauss_pdf(a[k] - b[k], sd_fac * sqrt(var[k * M->n + k])) / der;
            if (isnan(p) || isinf(p) || p < 0.0) return 1;
            Lp += log(p);
        } else {
            if (isnan(der) || isinf(der) || der <= 0.0) return 1;
            Lp -= log(der);
        }
    }
    if (isnan(Lp) || isinf(Lp)) return 1;
    *out = Lp;
    return 0;
}

/* eval_var_emp (mvn.c:196-233): one-pass covariance of the accepted thetas */
static void eval_var_emp(const mcmc_t *M, chain_t *c, double m)
{
    const int n = M->n;
    for (int i = 0; i < n; i++)
        for (int k = 0; k <= i; k++) {
            const double val = ((m - 1.0) / m) * (c->mean[i] - c->mean_sampling[i]) * (c->mean[k] - c->mean_sampling[k]);
            c->var_sampling[i * n + k] = ((m - 1.0) * c->var_sampling[i * n + k] + val) / m;
        }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < i; k++) c->var_sampling[k * n + i] = c->var_sampling[i * n + k];
    for (int i = 0; i < n; i++) c->mean_sampling[i] += (c->mean[i] - c->mean_sampling[i]) / m;
}

static void fmt_duration(double secs, char *buf, size_t len)
{
    int dd = (int)(secs / 86400.0), hh = (int)((secs - dd * 86400.0) / 3600.0);
    int mm = (int)((secs - dd * 86400.0 - hh * 3600.0) / 60.0);
    snprintf(buf, len, "%dd %dh %dm %gs", dd, hh, mm, secs - dd * 86400.0 - hh * 3600.0 - mm * 60.0);
}

static double elapsed(const struct timespec *t0)
{
    struct timespec t1;
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0->tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0->tv_nsec);
}

/* ------------------------------------------------------------------ main */
int main(int argc, char *argv[])
{
    int J = 1, nb_obs = -1, general_id = 0, n_threads = 1, fp32 = 0, device = 0, n_chains = 1, lanes = 0;
    int M_iter = 10, n_traj = 1000, print_acc = 0, pipe = 0, dry_run = 0;
    long seed_opt = -1;
    double dt = 0.0;
    unsigned int noises_off = 0;
    double eps_abs = 1e-6, eps_rel = 1e-3;   /* PLOM_EPS_ABS / PLOM_EPS_REL (C/core/plom.h:90-91) */
    char path[1024] = "./";
    mcmc_t MC;
    memset(&MC, 0, sizeof MC);
    MC.a = 0.999;
    MC.epsilon_max = 50.0;
    MC.alpha = 0.02;
    MC.m_switch = -1;
    MC.m_eps = 50;
    static int full_flag = 0, smooth_flag = 0;
    static struct option long_options[] = {
        {"full", no_argument, &full_flag, 1}, {"smooth", no_argument, &smooth_flag, 1},
        {"acc", no_argument, 0, 'r'}, {"no_dem_sto", no_argument, 0, 'x'},
        {"no_white_noise", no_argument, 0, 'y'}, {"no_diff", no_argument, 0, 'z'},
        {"DT", required_argument, 0, 's'}, {"eps_abs", required_argument, 0, 'v'},
        {"eps_rel", required_argument, 0, 'w'}, {"freeze_forcing", required_argument, 0, 'g'},
        {"n_traj", required_argument, 0, 'n'}, {"switch", required_argument, 0, 'S'},
        {"epsilon", required_argument, 0, 'E'}, {"cooling", required_argument, 0, 'a'},
        {"epsilon_max", required_argument, 0, 'f'}, {"alpha", required_argument, 0, 'G'},
        {"help", no_argument, 0, 'h'}, {"path", required_argument, 0, 'p'},
        {"id", required_argument, 0, 'i'}, {"n_thread", required_argument, 0, 'N'},
        {"LIKE_MIN", required_argument, 0, 'l'}, {"iter", required_argument, 0, 'M'},
        {"zmq", no_argument, 0, 'Z'}, {"chunk", required_argument, 0, 'c'},
        {"nb_obs", required_argument, 0, 'o'}, {"quiet", no_argument, 0, 'q'},
        {"pipe", no_argument, 0, 'P'}, {"chains", required_argument, 0, 1000},
        {"lanes", required_argument, 0, 1001}, {"fp32", no_argument, 0, 1002},
        {"device", required_argument, 0, 1003}, {"seed", required_argument, 0, 1004},
        {"trace_17g", no_argument, 0, 1005}, {"dry_run", no_argument, 0, 1006}, {0, 0, 0, 0}};
    const char *help =
        "PLOM pMCMC (MI355X)\n"
        "usage:\n"
        "pmcmc [implementation] [--no_dem_sto] [--no_white_noise] [--no_diff]\n"
        "                [-s, --DT <float || 0.25 day>]\n"
        "                [--full] [-n, --n_traj <int || 1000>] [--acc] [-p, --path <path>] [-i, --id <integer || 0>]\n"
        "                [-l, --LIKE_MIN <float || 1e-17>] [-J <integer || 1>] [-M, --iter <integer || 10>]\n"
        "                [-a --cooling <float || 0.999>] [-S --switch <int || 5*n_par_fitted^2 >]\n"
        "                [-E --epsilon <int || 50>] [--epsilon_max <float || 50.0>] [--smooth] [--alpha <float || 0.02>]\n"
        "                [-o, --nb_obs <int>] [-q, --quiet] [-P, --pipe]\n"
        "                [--chains <K || 1>] [--lanes <k>] [--fp32] [--device <k>] [--seed <k>] [--trace_17g] [-h, --help]\n"
        "where implementation is 'ode', 'sde' or 'psr' (default)\n";
    for (;;) {
        int idx = 0;
        int ch = getopt_long(argc, argv, "qPhrxyzs:v:w:i:J:l:M:p:c:N:ZS:E:a:f:G:n:o:g:", long_options, &idx);
        if (ch == -1) break;
        switch (ch) {
        case 0: break;
        case 'x': noises_off |= PLOM_NO_DEM_STO_BIT; break;
        case 'y': noises_off |= PLOM_NO_ENV_STO_BIT; break;
        case 'z': noises_off |= PLOM_NO_DRIFT_BIT; break;
        case 's': dt = atof(optarg); break;
        case 'v': eps_abs = atof(optarg); break;
        case 'w': eps_rel = atof(optarg); break;
        case 'c': case 'Z': break;
        case 'g': CLI_FREEZE = atof(optarg); break;
        case 'o': nb_obs = atoi(optarg); break;
        case 'a': MC.a = atof(optarg); break;
        case 'S': MC.m_switch = atoi(optarg); break;
        case 'E': MC.m_eps = atoi(optarg); break;
        case 'f': MC.epsilon_max = atof(optarg); break;
        case 'G': MC.alpha = atof(optarg); break;
        case 'h': printf("%s", help); return 1;
        case 'p': snprintf(path, sizeof path, "%s", optarg); break;
        case 'N': n_threads = atoi(optarg); break;
        case 'i': general_id = atoi(optarg); break;
        case 'J': J = atoi(optarg); break;
        case 'l': LIKE_MIN = atof(optarg); break;
        case 'M': M_iter = atoi(optarg); break;
        case 'r': print_acc = 1; break;
        case 'n': n_traj = atoi(optarg); break;
        case 'q': CLI_QUIET = 1; break;
        case 'P': pipe = 1; CLI_QUIET = 1; break;
        case 1000: n_chains = atoi(optarg); break;
        case 1001: lanes = atoi(optarg); break;
        case 1002: fp32 = 1; break;
        case 1003: device = atoi(optarg); break;
        case 1004: seed_opt = atol(optarg); break;
        case 1005: CLI_TRACE_17G = 1; break;
        case 1006: dry_run = 1; break;
        default: return 1;
        }
    }
    MC.full = full_flag;
    MC.smooth = smooth_flag;
    argc -= optind;
    argv += optind;
    const int impl = cli_implementation(argc, argv, help);
    uint32_t impl_flags = 0;
    cli_apply_implementation(impl, fp32, &noises_off, &impl_flags);
    const int psr = impl == CLI_PSR;
    if (n_chains < 1 || n_chains > 4096) cli_die("--chains must be in [1, 4096]");
    if (pipe && n_chains > 1) cli_die("-P (pipe) prints one theta: use it with one chain");
    if (M_iter < 1 || M_iter >= (1 << 30)) cli_die("-M must be in [1, 2^30)");
    if (J < 1) cli_die("-J must be at least 1");

    char str[4096];
    for (int k = 0; k < n_chains; k++) {
        snprintf(str, sizeof str, "%s.done_%d.json", path, general_id + k);
        unlink(str);
    }

    cli_problem P;
    cli_load(&P, 1, nb_obs, dt);
    P.psr = psr;
    P.drift_on = P.n_drift > 0 && !(noises_off & PLOM_NO_DRIFT_BIT) && !(impl_flags & PLOM_FLAG_ODE);
    MC.P = &P;
    const int n = P.n_theta, N = P.N, S = P.mi.size_proj, N_TS = P.N_TS;
    MC.n = n;

    /* build_best / load_best (build.c:1504-1601, load.c:304-406): theta, prior, initial
     * covariance (sd_transf^2 on the diagonal, or theta's "covariance"), followers and the
     * theta remainder are not estimated; white-noise sd parameters are not with
     * --no_white_noise */
    double *theta0 = calloc((size_t)n + 1, sizeof(double));
    cli_theta_from_guess(&P, theta0);
    MC.var0 = calloc((size_t)n * n + 1, sizeof(double));
    MC.is_estimated = calloc((size_t)n + 1, sizeof(int));
    MC.tbe0 = calloc((size_t)n + 1, sizeof(int));
    MC.tr = calloc((size_t)n + 1, sizeof(cli_tr));
    MC.pmin = calloc((size_t)n + 1, sizeof(double));
    MC.pmax = calloc((size_t)n + 1, sizeof(double));
    MC.prior = calloc((size_t)n + 1, sizeof(cli_prior));
    int *is_follower = calloc((size_t)n + 1, sizeof(int));
    for (int i = 0; i < P.n_par; i++)
        for (int g = 0; g < P.R[i].n_gp; g++) {
            const int o = P.offset[i] + g;
            MC.tr[o] = P.R[i].tr[g];
            MC.pmin[o] = P.R[i].min[g];
            MC.pmax[o] = P.R[i].max[g];
            MC.prior[o] = P.R[i].prior[g];
            is_follower[o] = P.R[i].follow >= 0;
            MC.var0[o * n + o] = (P.R[i].follow >= 0 || P.R[i].is_remainder) ? 0.0 : P.R[i].sd[g];
        }
    if (!cli_load_covariance(&P, MC.var0)) {
        for (int o = 0; o < n; o++)
            if (MC.var0[o * n + o] > 0.0) MC.var0[o * n + o] *= MC.var0[o * n + o];
    }
    if (noises_off & PLOM_NO_ENV_STO_BIT)
        for (int i = 0; i < P.n_par; i++)
            if (P.noise_par[i])
                for (int g = 0; g < P.R[i].n_gp; g++) {
                    const int o = P.offset[i] + g;
                    for (int k = 0; k < n; k++) MC.var0[o * n + k] = MC.var0[k * n + o] = 0.0;
                }
    /* update_to_be_estimated (func.c:171-187) */
    for (int o = 0; o < n; o++) {
        if (MC.var0[o * n + o] > 0.0 && !is_follower[o]) {
            MC.tbe0[MC.n_tbe++] = o;
            MC.is_estimated[o] = 1;
        } else {
            MC.var0[o * n + o] = 0.0;
        }
    }
    if (MC.full && MC.n_tbe == 0) cli_die("--full needs at least one parameter with sd_transf > 0");
    const int min_switch = 5 * MC.n_tbe * MC.n_tbe;
    if (MC.m_switch < 0) {
        MC.m_switch = min_switch;
    } else if (MC.m_switch < min_switch) {
        snprintf(str, sizeof str, "attention: covariance switching iteration (%i) is smaller than proposed one (%i)\n",
                 MC.m_switch, min_switch);
        fprintf(stderr, "%s", str);
    }
    if (dry_run) {
        /* the chain's inputs as JSON, no GPU call (for tests of the host-side logic) */
        double *pv = calloc((size_t)P.mi.pv_size, sizeof(double)), *x0 = calloc((size_t)S, sizeof(double));
        cli_pack(&P, theta0, pv, x0);
        static const char *trn[] = {"identity", "log", "logit", "logit_ab"};
        static const char *prn[] = {"flat", "normal", "pseudo_uniform"};
        printf("{\"model\": \"%s\", \"N\": %d, \"dt\": %.17g, \"m_switch\": %d, \"theta\": [", P.model_name, N, P.dt,
               MC.m_switch);
        for (int o = 0; o < n; o++) printf("%s%.17g", o ? ", " : "", theta0[o]);
        printf("], \"var\": [");
        for (int o = 0; o < n * n; o++) printf("%s%.17g", o ? ", " : "", MC.var0[o]);
        printf("], \"estimated\": [");
        for (int o = 0; o < MC.n_tbe; o++) printf("%s%d", o ? ", " : "", MC.tbe0[o]);
        printf("], \"components\": [");
        for (int i = 0, first = 1; i < P.n_par; i++)
            for (int g = 0; g < P.R[i].n_gp; g++, first = 0) {
                const int o = P.offset[i] + g;
                printf("%s{\"name\": \"%s:%s\", \"tr\": \"%s\", \"min\": %.17g, \"max\": %.17g, \"prior\": \"%s\", "
                       "\"follow\": %d, \"remainder\": %d}",
                       first ? "" : ", ", P.R[i].name, P.R[i].group_name[g], trn[MC.tr[o]], MC.pmin[o], MC.pmax[o], prn[MC.prior[o]],
                       P.R[i].follow, P.R[i].is_remainder);
            }
        printf("], \"pv\": [");
        for (int k = 0; k < P.mi.pv_size; k++) printf("%s%.17g", k ? ", " : "", pv[k]);
        printf("], \"x0\": [");
        for (int k = 0; k < S; k++) printf("%s%.17g", k ? ", " : "", x0[k]);
        printf("]}\n");
        return 0;
    }
    n_traj = n_traj < M_iter ? n_traj : M_iter;
    const int thin_traj = n_traj > 0 ? (int)((double)M_iter / (double)n_traj) : 0;
    const int print_x = n_traj > 0 && N > 0;

    /* chains */
    const uint32_t seed0 = (uint32_t)(seed_opt >= 0 ? seed_opt : 2 + general_id);
    chain_t *C = calloc((size_t)n_chains, sizeof(chain_t));
    for (int k = 0; k < n_chains; k++) {
        chain_t *c = &C[k];
        c->id = general_id + k;
        c->seed = seed0 + (uint32_t)k;
        c->mean = calloc((size_t)n + 1, sizeof(double));
        c->proposed = calloc((size_t)n + 1, sizeof(double));
        c->mean_sampling = calloc((size_t)n + 1, sizeof(double));
        c->var_sampling = calloc((size_t)n * n + 1, sizeof(double));
        c->tbe = calloc((size_t)MC.n_tbe + 1, sizeof(int));
        c->ar = calloc((size_t)n + 1, sizeof(double));
        c->sm_ar = calloc((size_t)n + 1, sizeof(double));
        memcpy(c->mean, theta0, (size_t)n * sizeof(double));
        memcpy(c->proposed, theta0, (size_t)n * sizeof(double));
        memcpy(c->tbe, MC.tbe0, (size_t)MC.n_tbe * sizeof(int));
        for (int o = 0; o < n; o++) c->ar[o] = c->sm_ar[o] = 1.0;
        /* build_mcmc_calc_data (mcmc_util.c:5-57) */
        c->has_cycled = 1;
        c->m_full = 0;
        c->cycle_id = MC.n_tbe - 1;
        c->epsilon = 1.0;
        c->pv = calloc((size_t)P.mi.pv_size, sizeof(double));
        c->x0 = calloc((size_t)S, sizeof(double));
        cli_pack(&P, c->mean, c->pv, c->x0);
        c->ftrace = cli_fopen(path, "trace", c->id);
        cli_trace_header(&P, c->ftrace);
        if (print_x) {
            c->fx = cli_fopen(path, "X", c->id);
            cli_x_header(&P, c->fx);
        }
        if (print_acc) {
            c->facc = cli_fopen(path, "acc", c->id);
            fprintf(c->facc, "index,epsilon,global_ar");
            if (!MC.full)
                for (int i = 0; i < P.n_par; i++)
                    for (int g = 0; g < P.R[i].n_gp; g++) fprintf(c->facc, ",%s:%s", P.R[i].name, P.R[i].group_name[g]);
            fprintf(c->facc, "\n");
        }
    }

    snprintf(str, sizeof str, "Starting plom-pmcmc with the following options: i = %d, J = %d, LIKE_MIN = %g, M = %d, n_threads = %d",
             general_id, J, LIKE_MIN, M_iter, n_threads);
    cli_log(str);
    if (n_chains > 1) {
        snprintf(str, sizeof str, "%d chains (ids %d..%d) evaluated as one batch per iteration", n_chains, general_id,
                 general_id + n_chains - 1);
        cli_log(str);
    }
    struct timespec t_all;
    clock_gettime(CLOCK_MONOTONIC, &t_all);

    /* the filters: one per chain, advanced together by one batch plan */
    plom_ctx *ctx = NULL;
    if (plom_ctx_create(device, &ctx) != PLOM_OK) cli_die(plom_last_error());
    plom_filter **fs = calloc((size_t)n_chains, sizeof(plom_filter *));
    for (int k = 0; k < n_chains; k++) {
        plom_filter_desc d;
        memset(&d, 0, sizeof d);
        d.model_id = P.model_id;
        d.J = J;
        d.N = N;
        d.precision = fp32 ? PLOM_FP32 : PLOM_FP64;
        d.noises_off = noises_off;
        d.eps_abs = eps_abs;
        d.eps_rel = eps_rel;
        d.flags = (print_x ? PLOM_FLAG_LINEAGE : 0u) | impl_flags;
        d.seed = C[k].seed;
        d.stream = 0;
        d.dt = P.dt;
        d.like_min = LIKE_MIN;
        plom_filter_inputs in = {P.times, P.data, C[k].pv, C[k].x0, NULL, P.has_cov ? P.cov : NULL,
                                  P.drift_on ? &P.dmap : NULL};
        if (plom_filter_create(ctx, &d, &in, &fs[k]) != PLOM_OK) cli_die(plom_last_error());
        C[k].f = fs[k];
    }
    plom_batch *batch = NULL;
    if (plom_batch_create(ctx, n_chains, fs, &batch) != PLOM_OK) cli_die(plom_last_error());
    if (lanes <= 0) lanes = n_chains < 4 ? n_chains : 4;
    if (plom_batch_configure(ctx, batch, lanes) != PLOM_OK) cli_die(plom_last_error());

    double *lin = print_x ? calloc((size_t)(N + 1) * S, sizeof(double)) : NULL;
    double *lob = print_x ? calloc((size_t)(N + 1) * N_TS * 2 + 1, sizeof(double)) : NULL;
    const int nt = MC.n_tbe;
    double *z = calloc((size_t)n + 4, sizeof(double));
    double *sub = calloc((size_t)nt * nt + 1, sizeof(double));
    double **Lk = calloc((size_t)n_chains, sizeof(double *));
    const double **vark = calloc((size_t)n_chains, sizeof(double *));
    double *sdk = calloc((size_t)n_chains, sizeof(double));
    for (int k = 0; k < n_chains; k++) Lk[k] = calloc((size_t)nt * nt + 1, sizeof(double));
    int warned_pd = 0;

    for (int m = 0; m < M_iter; m++) {
        for (int k = 0; k < n_chains; k++) clock_gettime(CLOCK_MONOTONIC, &C[k].t_begin);
        /* proposals */
        if (m > 0) {
            for (int k = 0; k < n_chains; k++) {
                chain_t *c = &C[k];
                /* increment_iteration_counters */
                if (MC.full) {
                    c->has_cycled = 1;
                    c->m_full++;
                    c->cycle_id = nt;
                } else if (c->cycle_id >= nt - 1) {
                    c->has_cycled = 1;
                    c->m_full++;
                    c->cycle_id = 0;
                } else {
                    c->has_cycled = 0;
                    c->cycle_id++;
                }
                /* get_var_and_sd_fac */
                const double *var = MC.var0;
                double sd_fac = 1.0;
                if (MC.full) {
                    if (m > MC.m_eps && m * c->global_ar < MC.m_switch) {
                        const double ar = MC.smooth ? c->sm_global_ar : c->global_ar;
                        c->epsilon *= exp(pow(MC.a, (double)(m - 1)) * (ar - 0.234));
                    } else {
                        c->epsilon = 1.0;
                    }
                    c->epsilon = c->epsilon < MC.epsilon_max ? c->epsilon : MC.epsilon_max;
                    sd_fac = c->epsilon * 2.38 / sqrt((double)nt);
                    if (m * c->global_ar >= MC.m_switch) var = c->var_sampling;
                    for (int i = 0; i < nt; i++)
                        for (int q = 0; q < nt; q++) sub[i * nt + q] = var[MC.tbe0[i] * n + MC.tbe0[q]];
                    if (cli_cholesky(sub, nt, sd_fac * sd_fac, Lk[k])) {
                        if (!warned_pd) fprintf(stderr, "Covariance matrix is not positive definite: the initial covariance is used\n");
                        warned_pd = 1;
                        var = MC.var0;
                        for (int i = 0; i < nt; i++)
                            for (int q = 0; q < nt; q++) sub[i * nt + q] = var[MC.tbe0[i] * n + MC.tbe0[q]];
                        if (cli_cholesky(sub, nt, sd_fac * sd_fac, Lk[k])) cli_die("the initial covariance is not positive definite");
                    }
                } else if (nt > 0 && c->has_cycled) {
                    /* gsl_ran_shuffle: Fisher-Yates, j uniform in [0, i] */
                    for (int i = nt - 1; i >= 1; i--) {
                        uint32_t w[4];
                        chain_words(c, (uint32_t)m, PURPOSE_SHUFFLE, (uint32_t)(i / 4), 0u, w);
                        const int j = (int)(((uint64_t)w[i % 4] * (uint64_t)(i + 1)) >> 32);
                        const int t = c->tbe[i];
                        c->tbe[i] = c->tbe[j];
                        c->tbe[j] = t;
                    }
                }
                vark[k] = var;
                sdk[k] = sd_fac;
                /* propose_safe_theta_and_load_X0 */
                int attempt = 0;
                for (;; attempt++) {
                    if (attempt > 1000000) cli_die("no proposal satisfies the initial-condition constraint");
                    if (MC.full) {
                        chain_normals(c, m, attempt, nt, z);
                        memcpy(c->proposed, c->mean, (size_t)n * sizeof(double));
                        for (int i = 0; i < nt; i++) {
                            double v = 0.0;
                            for (int q = 0; q <= i; q++) v += Lk[k][i * nt + q] * z[q];
                            c->proposed[MC.tbe0[i]] = c->mean[MC.tbe0[i]] + v;
                        }
                    } else if (nt > 0) {
                        chain_normals(c, m, attempt, 1, z);
                        const int o = c->tbe[c->cycle_id];
                        c->proposed[o] = c->mean[o] + sd_fac * sqrt(var[o * n + o]) * z[0];
                    }
                    if (cli_theta_remainder(&P, c->proposed) == 0) break;
                }
                cli_follow(&P, c->proposed);
                cli_pack(&P, c->proposed, c->pv, c->x0);
                if (plom_filter_update(ctx, c->f, c->pv, c->x0, c->seed, (uint32_t)m) != PLOM_OK) cli_die(plom_last_error());
            }
        }
        /* one batched pass: every chain's filter */
        if (plom_batch_launch(ctx, batch, 1) != PLOM_OK) cli_die(plom_last_error());
        const int x_now = print_x;
        for (int k = 0; k < n_chains; k++) {
            chain_t *c = &C[k];
            plom_filter_outputs out;
            memset(&out, 0, sizeof out);
            int want_x = 0;
            if (x_now) {
                if (m == 0) want_x = 1;
                else if ((MC.full || c->cycle_id >= nt - 1) && thin_traj > 0 && (c->m_full % thin_traj) == 0) want_x = 1;
            }
            if (want_x) {
                out.lineage = lin;
                out.lineage_obs = lob;
            }
            if (plom_filter_read(ctx, c->f, &out) != PLOM_OK) cli_die(plom_last_error());
            c->L_new = out.loglik;
            char pre[64] = "";
            if (n_chains > 1) snprintf(pre, sizeof pre, "chain %d: ", c->id);
            char dur[128];
            if (m == 0) {
                c->L_prev = c->L_new;
                if (want_x) cli_x_rows(&P, c->fx, 0, lin, lob);
                cli_trace_row(&P, c->ftrace, 0, c->mean, c->L_new);
                fmt_duration(elapsed(&c->t_begin), dur, sizeof dur);
                snprintf(str, sizeof str, "%siteration number:%d\t logV: %g\t accepted:%d computed in:= %s", pre, m, c->L_new,
                         c->accept, dur);
                cli_log(str);
                continue;
            }
            /* metropolis_hastings */
            double Lqn = 0, Lqp = 0, Lpn = 0, Lpp = 0, alpha = 0.0;
            const double *Lf = MC.full ? Lk[k] : NULL;
            const int rc1 = log_prob_proposal(&MC, c->proposed, c->mean, vark[k], sdk[k], Lf, &Lqn);
            const int rc2 = log_prob_proposal(&MC, c->mean, c->proposed, vark[k], sdk[k], Lf, &Lqp);
            const int rc3 = log_prob_prior(&MC, c->proposed, &Lpn);
            const int rc4 = log_prob_prior(&MC, c->mean, &Lpp);
            int accepted = 0;
            if (!rc1 && !rc2 && !rc3 && !rc4) {
                alpha = exp(c->L_new - c->L_prev + Lqp - Lqn + Lpn - Lpp);
                uint32_t w[4];
                chain_words(c, (uint32_t)m, PURPOSE_ACCEPT, 0u, 0u, w);
                const double u = ((double)w[0] + 0.5) * 0x1p-32;
                accepted = u < alpha;
            }
            if (accepted) {
                c->L_prev = c->L_new;
                memcpy(c->mean, c->proposed, (size_t)n * sizeof(double));
                c->accept++;
            } else if (!MC.full && nt > 0) {
                const int o = c->tbe[c->cycle_id];
                c->proposed[o] = c->mean[o];
            }
            /* compute_acceptance_rates */
            const double acc = (double)accepted;
            if (!MC.full && nt > 0) {
                const int o = c->tbe[c->cycle_id];
                c->sm_ar[o] = (1.0 - MC.alpha) * c->sm_ar[o] + MC.alpha * acc;
                c->ar[o] += (acc - c->ar[o]) / (double)c->m_full;
            }
            c->sm_global_ar = (1.0 - MC.alpha) * c->sm_global_ar + MC.alpha * acc;
            c->global_ar += (acc - c->global_ar) / (double)m;
            fmt_duration(elapsed(&c->t_begin), dur, sizeof dur);
            snprintf(str, sizeof str, "%siteration number: %d (%d / %d)\t logV: %g (previous was %g) accepted: %d computed in:= %s",
                     pre, c->m_full, c->cycle_id, nt, c->L_new, c->L_prev, c->accept, dur);
            cli_log(str);
            if (c->cycle_id >= nt - 1) {
                eval_var_emp(&MC, c, (double)c->m_full);
                if (want_x) cli_x_rows(&P, c->fx, c->m_full, lin, lob);
                cli_trace_row(&P, c->ftrace, c->m_full, c->mean, c->L_prev);
                if (c->facc) {
                    fprintf(c->facc, "%d,%g,%g", c->m_full, c->epsilon, c->global_ar);
                    if (!MC.full)
                        for (int o = 0; o < n; o++) fprintf(c->facc, ",%g", c->ar[o]);
                    fprintf(c->facc, "\n");
                }
                snprintf(str, sizeof str, "%sacceptance rate(s) at iteration %d: %g (smoothed: %g)", pre, c->m_full, c->global_ar,
                         c->sm_global_ar);
                cli_log(str);
            }
        }
    }

    /* outputs of the end: covariance_<id>.csv (header_covariance / print_covariance) and the theta */
    for (int k = 0; k < n_chains; k++) {
        chain_t *c = &C[k];
        fclose(c->ftrace);
        if (c->fx) fclose(c->fx);
        if (c->facc) fclose(c->facc);
        FILE *fc = cli_fopen(path, "covariance", c->id);
        int first = 1;
        for (int i = 0; i < P.n_par; i++)
            if (i != P.ind_theta_remainder)
                for (int g = 0; g < P.R[i].n_gp; g++) {
                    fprintf(fc, "%s%s:%s", first ? "" : ",", P.R[i].name, P.R[i].group_name[g]);
                    first = 0;
                }
        fprintf(fc, "\n");
        for (int i = 0; i < P.n_par; i++) {
            if (i == P.ind_theta_remainder) continue;
            for (int g = 0; g < P.R[i].n_gp; g++) {
                const int row = P.offset[i] + g;
                first = 1;
                for (int ii = 0; ii < P.n_par; ii++) {
                    if (ii == P.ind_theta_remainder) continue;
                    for (int gg = 0; gg < P.R[ii].n_gp; gg++) {
                        fprintf(fc, "%s%g", first ? "" : ",", c->var_sampling[row * n + P.offset[ii] + gg]);
                        first = 0;
                    }
                }
                fprintf(fc, "\n");
            }
        }
        fclose(fc);
        cli_done(&P, c->mean, path, c->id, pipe);
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
