/*
 * plom_smc.h -- C ABI of the MI355X (gfx950) PLOM sequential Monte Carlo filter.
 *
 * Plain C: no torch or HIP types in any signature; the caller owns every host buffer,
 * the library owns every device buffer.  All functions return PLOM_OK (0) or a negative
 * PLOM_ERR_* code and never exit(); plom_last_error() gives the message of the last
 * failure on the calling thread.  One context per host thread (contexts are not locked).
 *
 * What each entry point replaces in the reference (paths relative to the reference
 * checkout, model_builder/C/ abbreviated C/):
 *
 *   plom_smc_run            run_SMC_zmq_inproc, C/core/smc.c:379-467 (and its OpenMP
 *                           twin run_SMC, :204-307) together with the per-thread worker
 *                           C/core/worker_inproc.c:21-102: propagate (f_prediction_sde_*,
 *                           C/core/prediction_util.c:197-287, generated step_sde_*,
 *                           C/templates/core_template.c:215-339), weight (get_log_likelihood,
 *                           C/core/likelihood.c:56-70; weight, smc.c:27-83), resample
 *                           (systematic_sampling, smc.c:90-112; resample_X, smc.c:142-161).
 *                           Called where C/smc/main.c:318 and pmcmc's run_propag
 *                           (C/pmcmc/pmcmc.c:24-37) call the filter.
 *   plom_filter_create/...  the same filter split into upload / run / read so that a
 *                           caller (pMCMC proposals, a server) keeps clouds resident.
 *   plom_batch_run          no counterpart: many independent filters (pMCMC proposals,
 *                           C/pmcmc/pmcmc.c:186-266 runs them one after another) advanced
 *                           together, one observation interval per batched launch.
 *   plom_server_*           no counterpart: continuous-batching request loop with
 *                           admission by free HBM.
 *   plom_model_*            the generated per-model API (C/core/plom.h:896-913,
 *                           C/templates/core_template.c) compiled into the library.
 *   PLOM_FLAG_MIF,          one pass of iterated filtering (C/mif/mif.c:128-200): every
 *   plom_filter_mif_start   particle carries its own parameter vector theta, resampled with
 *                           its state and moved by a random walk after every observation
 *                           (resample_and_mut_theta_mif, C/mif/methods.c:261-367); the pass
 *                           returns the weighted mean and the variance of theta per
 *                           observation (mean_var_theta_theoretical_mif, methods.c:146-205).
 */
#ifndef PLOM_SMC_H
#define PLOM_SMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PLOM_OK 0
#define PLOM_ERR_ARG -1      /* invalid argument or shape */
#define PLOM_ERR_HIP -2      /* HIP runtime error */
#define PLOM_ERR_NOMEM -3    /* device memory exhausted */
#define PLOM_ERR_MODEL -4    /* unknown model id */
#define PLOM_ERR_STATE -5    /* call out of order (e.g. read before run) */

/* noise switches, same bits as the reference's enum plom_noises_off (C/core/plom.h:66) */
#define PLOM_NO_DEM_STO_BIT 1u
#define PLOM_NO_ENV_STO_BIT 2u
#define PLOM_NO_DRIFT_BIT 4u

#define PLOM_FP64 0
#define PLOM_FP32 1 /* FP32 propagation; log-weights, sums, scan and log-likelihood stay FP64 */

#define PLOM_FLAG_NO_FILTER 1u        /* -t/--no_filter: propagate only (C/smc/main.c:177) */
#define PLOM_FLAG_STORE_ANCESTRY 2u   /* keep select[n][j] for every n (needed by --full) */
#define PLOM_FLAG_FINAL_STATE 4u      /* keep the resampled final cloud */
#define PLOM_FLAG_HAT 16u             /* per interval: weighted mean and 95 % bounds of every state,
                                         remainder and observation mean (compute_hat, C/core/hat.c:66-137)
                                         and the prediction residuals (print.c:632-701) */
#define PLOM_FLAG_LINEAGE 8u          /* sample one particle of the last interval by its weight and
                                         return its whole path (pmcmc --full,
                                         sample_traj_and_print, C/core/print.c:715-784); keeps the
                                         ancestry on the device */

#define PLOM_FLAG_PSR 64u             /* implementation "psr" (the reference's default): the Poisson
                                         system with stochastic rates, an Euler-multinomial step with
                                         gamma white noise (step_psr, C/templates/core_template.c:104-209);
                                         FP64 only; of noises_off only PLOM_NO_ENV_STO_BIT applies */
#define PLOM_FLAG_ODE 128u            /* implementation "ode": the rate equations integrated without noise by
                                         adaptive Runge-Kutta-Fehlberg 4(5) under the error control
                                         eps_abs / eps_rel (f_prediction_ode, C/core/prediction_util.c:176-193);
                                         FP64 only; noises_off is ignored; time-varying covariates are
                                         evaluated as splines at the integrator's times */
#define PLOM_FLAG_ALL_STATES 256u     /* keep every particle's state after every interval (smc -r, print_X,
                                         C/core/print.c:360-392): plom_filter_outputs.all_states /
                                         all_obs; keeps the ancestry */
#define PLOM_FLAG_MIF 32u             /* per-particle parameters (iterated filtering): needs
                                         plom_filter_inputs.mif and plom_filter_mif_start before
                                         every run; not with NO_FILTER, HAT or LINEAGE */

#define PLOM_MAX_J (1 << 21)          /* particles per filter */

/* school terms (terms_forcing, C/core/special_functions.c:21-47): holiday periods per
 * population, and the values of one population's record in pv */
#ifndef PLOM_MAX_TERMS
#define PLOM_MAX_TERMS 8
#define PLOM_SCHOOL_REC (2 + 2 * PLOM_MAX_TERMS)
#endif

typedef struct {
    int32_t size_proj;   /* S: N_PAR_SV*N_CAC + N_TS_INC_UNIQUE (C/smc/main.c:249) */
    int32_t n_ts;        /* N_TS */
    int32_t n_cac;       /* N_CAC */
    int32_t n_par_sv;    /* N_PAR_SV */
    int32_t inc_offset;  /* first incidence slot of the state */
    int32_t n_inc;       /* N_TS_INC_UNIQUE */
    int32_t pv_size;     /* length of the flat natural-parameter vector; a model whose rates use
                            terms_forcing ends it with one school record per population
                            (PLOM_SCHOOL_REC values: n_terms, prop_school, PLOM_MAX_TERMS
                            (begin, end) holiday periods in years) */
    int32_t n_dem;       /* demographic noises per population */
    int32_t n_env;       /* environmental noises per population */
    int32_t n_hat;       /* quantities of a hat row: states, remainders, observation means */
    int32_t n_pp;        /* leading slots of pv that PLOM_FLAG_MIF makes per particle: the process
                            parameters [par][n_cac] then the observation parameters [par][n_ts] */
    int32_t fix_off;     /* first slot of pv holding a covariate (metadata: population size, birth
                            rate, proportion ...): [covariate][n_cac] then those of the observation
                            model [covariate][n_ts] */
    int32_t n_fix;       /* number of those slots (plom_filter_inputs.cov has n_fix series) */
    int32_t n_drift;     /* drifting process parameters (diffusions); their states sit at drift_offset,
                            one per parameter and population: [n_drift][n_cac] */
    int32_t drift_offset;
} plom_model_info;

typedef struct {
    int32_t model_id;
    int32_t J;           /* particles, 1..PLOM_MAX_J */
    int32_t N;           /* observation intervals (nb_obs) */
    int32_t precision;   /* PLOM_FP64 or PLOM_FP32 */
    uint32_t noises_off; /* PLOM_NO_*_BIT */
    uint32_t flags;      /* PLOM_FLAG_* */
    uint32_t seed;       /* random key, first word */
    uint32_t stream;     /* random key, second word: one per filter / request */
    double dt;           /* Euler step, already 1/round(1/DT) (C/core/build.c:1298-1304) */
    double like_min;     /* LIKE_MIN (C/smc/main.c:83) */
    int32_t n_theta;     /* PLOM_FLAG_MIF: components of the per-particle parameter vector theta */
    int32_t pad;
    double eps_abs;      /* PLOM_FLAG_ODE: absolute and relative error targets of the integration; */
    double eps_rel;      /* both 0: the reference's defaults 1e-6 and 1e-3 (C/core/plom.h:90-91) */
} plom_filter_desc;

/* How a particle's theta (transformed scale, one value per parameter and group, as
 * p_best->mean) gives the per-particle slots of pv (back_transform_theta2par,
 * C/core/transform.c:462-514) and how it moves (PLOM_FLAG_MIF).  Arrays over the
 * components have n_theta entries. */
typedef struct {
    const int32_t *pv_src;     /* [n_pp]: the component whose natural value fills the slot, or -1
                                  (the slot keeps pv) */
    const int32_t *tr;         /* 0 identity, 1 log, 2 logit, 3 logit_ab (on [min, max]) */
    const int32_t *scale;      /* 0 none, 1 10^x, 2 10^-x (applied after the inverse transform) */
    const int32_t *duration;   /* 1: the value is a duration, the rate is 1/(1e-5 + value) */
    const double *min, *max;   /* bounds of logit_ab */
    const double *multiplier;  /* unit multiplier */
    const double *sd;          /* random-walk standard deviation per unit of mut_fac (0: the
                                  component is resampled but not moved) */
    const double *chol;        /* optional [n_theta * n_theta], n_theta <= 64: a lower-triangular factor L
                                  (row-major) of the random walk's covariance per unit of mut_fac^2
                                  (a "covariance" in theta, resample_and_mut_theta_mif with is_mvn,
                                  C/mif/methods.c:275-330): theta_o moves by mut_fac sum_{q<=o}
                                  L[o][q] z_q instead of mut_fac sd_o z_o; rows of zeros do not move */
    /* optional (all three or none): the prior patch of the weights (mif --prior,
     * patch_likelihood_prior, C/mif/methods.c:469-523).  At an interval n with data, particle
     * j's log-weight gains sum over o with prior[o] >= 0 and n < prior_until[o] of
     * prior_pow[o] log(max(p_o, like_min)), p_o the density of f_inv(theta_o) (its natural
     * value before scale, multiplier and duration) under prior[o] on [min[o], max[o]]; the
     * particle fails when the patched weight is at or below like_min^n_ts.  The reference
     * uses prior_pow 1/nb_obs for the process / observation parameters at every interval and
     * 1/L for the initial conditions at intervals n < L. */
    const int32_t *prior;      /* [n_theta] PLOM_PRIOR_*: -1 none, 0 uniform, 1 normal, 2 pseudo-uniform */
    const double *prior_pow;   /* [n_theta] exponent of the density */
    const int32_t *prior_until;/* [n_theta] the term applies at intervals n < prior_until[o] */
} plom_mif_map;

#define PLOM_PRIOR_NONE (-1)
#define PLOM_PRIOR_UNIFORM 0
#define PLOM_PRIOR_NORMAL 1
#define PLOM_PRIOR_PSEUDO_UNIFORM 2

/* A covariate as a function of time, for one slot of pv (one covariate of one population or
 * series), as the reference's spline of par_fixed_values (build_p_calc,
 * C/core/build.c:1075-1211, evaluated by the generated code, model_builder/Ccoder.py:39-43):
 * n >= 4 knots: natural cubic spline; 2 or 3: linear; 1: constant.  The filter evaluates it at
 * the start of every Euler substep (rates) and at every observation time (observation model,
 * remainders); every such time must lie within [x[0], x[n-1]] when n > 1. */
typedef struct {
    int32_t n;
    int32_t pad;
    const double *x;     /* [n] strictly increasing, days since date_0 */
    const double *y;     /* [n] values, in the data's units */
} plom_covariate;

/* The drifting parameters of a model with diffusions (compute_drift, C/core/drift.c:27-42;
 * back-transform in step_sde_*, C/templates/core_template.c:251-271), per drifting
 * parameter i and population c at index i * n_cac + c.  The drift state of (i, c) lives in
 * x0 / the state at drift_offset + i * n_cac + c, on the parameter's transformed scale (its
 * theta value, theta_driftIC2Xdrift, C/core/prediction_util.c:119-131); populations of one
 * group start equal and move together.  After every Euler substep the state of group g
 * moves by (pv[vol_slot[i * n_cac + g]] sqrt(dt)) z with one standard normal z per group, in group
 * order; the rates read the natural value of the state. */
typedef struct {
    const int32_t *group;      /* [n_drift * n_cac] group of each population, 0 <= g < n_cac */
    const int32_t *tr;         /* [n_drift * n_cac] 0 identity, 1 log, 2 logit, 3 logit_ab */
    const int32_t *scale;      /* 0 none, 1 10^x, 2 10^-x */
    const int32_t *duration;   /* 1: a duration, the rate is 1/(1e-5 + value) */
    const double *min, *max;   /* bounds of logit_ab */
    const double *multiplier;  /* unit multiplier */
    const int32_t *vol_slot;   /* [n_drift * n_cac]: entry i * n_cac + g is the pv slot holding the natural
                                  volatility of group g (natural[vol][g]; -1: none), so that a new pv
                                  (plom_filter_update) moves the volatility too */
} plom_drift_map;

typedef struct {
    const double *times; /* [N+1] days since date_0 (settings data.times) */
    const double *data;  /* [N*n_ts] observations, NaN = missing */
    const double *pv;    /* [pv_size] natural parameters per population / series */
    const double *x0;    /* [size_proj] initial state (incidence slots ignored) */
    const plom_mif_map *mif; /* PLOM_FLAG_MIF only (copied at creation) */
    const plom_covariate *cov; /* optional [n_fix]: time-varying covariates, evaluated at creation
                                  (they then replace the covariate slots of pv); NULL: the
                                  covariates are the constants in pv; a served request's are
                                  copied at submission */
    const plom_drift_map *drift; /* models with diffusions, unless noises_off has PLOM_NO_DRIFT_BIT or
                                    the implementation is ode (then the parameters keep their pv
                                    values); with PLOM_FLAG_MIF the volatilities are the particle's
                                    own values of their pv slots */
} plom_filter_inputs;

typedef struct {
    double loglik;       /* sum_n loglik_n, accumulated in step order (smc.c:80) */
    int32_t n_all_fail;  /* steps where every particle failed (smc.c:55-57) */
    double *loglik_n;    /* optional [N]: per-interval log-likelihood (0 without data) */
    double *ess_n;       /* optional [N]: effective sample size (0 without data / all fail) */
    int32_t *ancestry;   /* optional [N*J]: select[n][j]; needs PLOM_FLAG_STORE_ANCESTRY */
    double *final_state; /* optional [size_proj*J] SoA resampled final cloud (PLOM_FLAG_FINAL_STATE) */
    double *final_logw;  /* optional [J] log-weights of the last interval, before resampling */
    double *lineage;     /* optional [(N+1)*size_proj]: the sampled path, row n = its state at
                            times[n] (row 0 = x0); needs PLOM_FLAG_LINEAGE */
    int32_t *lineage_slots; /* optional [N]: the particle index of that path at each interval */
    double *lineage_obs; /* optional [(N+1)*n_ts*2]: per row and series, the observation mean along
                            the path and one observation drawn from the observation model */
    double *hat;         /* optional [(N+1)*n_hat*3]: per row (times[0..N]) and quantity: low95,
                            mean, high95 (PLOM_FLAG_HAT) */
    double *pred_res;    /* optional [N*n_ts*2]: per interval and series: prediction mean and
                            residual over the resampled cloud; NaN without data (PLOM_FLAG_HAT) */
    double *mif_mean;    /* optional [(N+1)*n_theta] (PLOM_FLAG_MIF): row n+1 = the weighted mean of
                            the particles' theta at interval n, NaN without data; row 0 NaN */
    double *mif_var;     /* optional [(N+1)*n_theta]: the same rows, unweighted variance over the
                            particles (divisor J-1) */
    double *all_states;  /* optional [(N+1)*J*size_proj] (PLOM_FLAG_ALL_STATES): row n, particle j = the
                            cloud after interval n-1 as resampling leaves it (resample_X,
                            C/core/smc.c:142-161; row 0 = x0) */
    double *all_obs;     /* optional [(N+1)*J*n_ts*2]: per row, particle and series, the observation
                            mean of that state at times[n] and one observation drawn from the
                            observation model (obs_mean / observation, core_template.c:369-412) */
} plom_filter_outputs;

typedef struct plom_ctx plom_ctx;
typedef struct plom_filter plom_filter;

const char *plom_last_error(void);

int plom_ctx_create(int device, plom_ctx **out);
int plom_ctx_destroy(plom_ctx *ctx);

/* The generated per-model API (C/core/plom.h:896-913, C/templates/core_template.c): the
 * models compiled into the library, then the models registered with plom_model_load. */
int plom_model_count(void);
int plom_model_find(const char *name);                 /* model id, or PLOM_ERR_MODEL */
int plom_model_name(int model_id, char *buf, int len);
int plom_model_info_get(int model_id, plom_model_info *out);
/* Register a user's model from its model library (built by `build.py --model <dir>` or
 * `plom_build.py --lib`: the generated code of that model compiled on its own, the role of
 * the reference's per-model build, scripts/pmbuilder:50-105 -> Builder.py:283-351 ->
 * C/templates/Makefile:28-29).  *model_id is the id to put in plom_filter_desc; loading a
 * library of an already registered model name gives that id again.  PLOM_ERR_MODEL if the
 * library cannot be loaded or was compiled against another version of this library.  The
 * library stays loaded until the process ends. */
int plom_model_load(const char *path, int32_t *model_id);

/* the value of a covariate at the times t[0..nt) (the population size at t = 0 for the initial
 * state, pop_size_t0 of C/core/build.c:1179-1185, and at the observation times for the
 * remainders written by the commands); PLOM_ERR_ARG for a time outside its knots */
int plom_covariate_eval(const plom_covariate *cov, int nt, const double *t, double *out);

/* one filter, blocking: upload, run, read back, free */
int plom_smc_run(plom_ctx *ctx, const plom_filter_desc *desc, const plom_filter_inputs *in,
                 plom_filter_outputs *out);

/* plom_smc_run without device allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The filter is laid out in a fixed device scratch of
 * PLOM_NOALLOC_MAX_BYTES that the library holds on every GPU it runs on until the process
 * ends; it makes no hipMalloc / hipFree and no other device or page-locked allocation (the
 * host-side staging of the inputs uses the C heap).  A filter whose device layout
 * (plom_filter_device_bytes) exceeds the scratch gets PLOM_ERR_ARG, and a context on a
 * device past the 64th has no scratch (PLOM_ERR_ARG).  Results are bit-identical to
 * plom_smc_run.  Every call on one GPU uses that GPU's one scratch, which keeps the last
 * call's contents: calls that overlap in time on one GPU (two threads, two contexts) race,
 * so the caller serialises them. */
#define PLOM_NOALLOC_MAX_BYTES ((size_t)64 << 20)
int plom_smc_run_noalloc(plom_ctx *ctx, const plom_filter_desc *desc, const plom_filter_inputs *in,
                         plom_filter_outputs *out);

/* resident filter: upload once, run (asynchronously on the context stream), read */
int plom_filter_create(plom_ctx *ctx, const plom_filter_desc *desc, const plom_filter_inputs *in,
                       plom_filter **out);
int plom_filter_reset(plom_ctx *ctx, plom_filter *f);  /* back to interval 0, same inputs */
/* new natural parameters [pv_size] and initial state [size_proj] (either may be NULL: kept)
 * and a new random key, in place (no reallocation; batch plans holding the filter pick the
 * change up at their next launch, graphs included): one pMCMC proposal (pmcmc.c:205-214) */
int plom_filter_update(plom_ctx *ctx, plom_filter *f, const double *pv, const double *x0, uint32_t seed,
                       uint32_t stream);
/* PLOM_FLAG_MIF: the particles' starting parameters theta [n_theta*J] (component-major) and
 * initial states x0 [size_proj*J] (state-major), the random-walk factor mut_fac [N] applied
 * after each interval with data (FREEZE*sqrt(delta_t), C/mif/mif.c:183), and the random key
 * of the pass.  A pass consumes them: call before every run (or batch launch). */
int plom_filter_mif_start(plom_ctx *ctx, plom_filter *f, const double *theta, const double *x0,
                          const double *mut_fac, uint32_t seed, uint32_t stream);
int plom_filter_run(plom_ctx *ctx, plom_filter *f);    /* enqueue every interval */
int plom_filter_read(plom_ctx *ctx, plom_filter *f, plom_filter_outputs *out); /* waits */
int plom_filter_destroy(plom_ctx *ctx, plom_filter *f);
int64_t plom_filter_device_bytes(const plom_filter_desc *desc); /* bytes a filter holds on the device */

/* Many resident filters advanced together: one batched launch round per interval for
 * all of them (uneven models, J and N allowed); each filter's results are bit-identical
 * to plom_filter_run.  A batch plan keeps the descriptors and tile maps on the device;
 * with use_graph the whole launch sequence (reset, every interval, final resampling) is
 * captured once as a hipGraph and replayed by every later launch.  The filters' inputs
 * may not change size while a plan holds them. */
typedef struct plom_batch plom_batch;
int plom_batch_create(plom_ctx *ctx, int n_filters, plom_filter **filters, plom_batch **out);
int plom_batch_launch(plom_ctx *ctx, plom_batch *batch, int use_graph);  /* asynchronous */
int plom_batch_destroy(plom_ctx *ctx, plom_batch *batch);
/* Lanes: the plan's filters split round-robin (by descending N) into `lanes` chains, each
 * advanced on a stream of its own, forked from and joined back into the context stream
 * (also inside the captured graph), so that one chain fills the SIMDs another leaves idle.
 * Results do not depend on the number of lanes.  plom_batch_configure sets it (drops a
 * captured graph); plom_batch_autotune times `rounds` launch rounds spread over the pass
 * (the grids of intervals 0, maxN/rounds, 2 maxN/rounds, ...) with 1, 2, 4 and 8 lanes
 * (at most one per filter), keeps the fastest and reports it (`ms_out`: 4 timings, 0 for
 * a skipped candidate).  The next launch starts the filters from interval 0 again. */
#define PLOM_MAX_LANES 8
int plom_batch_configure(plom_ctx *ctx, plom_batch *batch, int lanes);
int plom_batch_autotune(plom_ctx *ctx, plom_batch *batch, int rounds, int32_t *lanes_out, double *ms_out);

/* Tiles per propagate block (no counterpart in the reference): a propagate block holds 1, 2
 * or 4 tiles of 256 particles (256, 512 or 1024 threads; every tile keeps its own weight
 * statistics, so results do not depend on it).  Larger blocks load the Gaussians' LDS table
 * once for more particles; smaller ones schedule more finely.  plom_batch_tune extends
 * plom_batch_autotune: after the lanes it tunes the tiles per block of every (model,
 * precision, variant, MIF) group of the batch in turn (1, 2, 4; the ode implementation and
 * MIF filters use 1) and reports every timing; plom_batch_autotune also does this and
 * reports the lanes part.  plom_ctx_set_tiles_per_block sets the value single filters and
 * the server use, and batch groups that were not tuned (default 1). */
#define PLOM_TUNE_MAX_GROUPS 16
typedef struct {
    int32_t model_id, precision, variant, mif;  /* variant: 0-3 sde noise switches, 4-5 psr, 6 ode */
    int32_t tiles_per_block;                     /* chosen */
    int32_t pad;
    double ms[3];                                /* the rounds' time with 1, 2, 4 tiles per block (0: not tried) */
} plom_tune_group;
typedef struct {
    int32_t rounds, lanes;                       /* launch rounds timed per candidate; lanes chosen */
    double lanes_ms[4];                          /* with 1, 2, 4, 8 lanes (0: not tried) */
    int32_t n_groups, pad;
    plom_tune_group group[PLOM_TUNE_MAX_GROUPS];
} plom_batch_tuning;
int plom_batch_tune(plom_ctx *ctx, plom_batch *batch, int rounds, plom_batch_tuning *out);
int plom_ctx_set_tiles_per_block(plom_ctx *ctx, int tiles_per_block);
int plom_batch_run(plom_ctx *ctx, int n_filters, plom_filter **filters, int use_graph); /* create + launch + destroy */

/* Served stream (no counterpart in the reference): continuous batching of filter
 * requests.  plom_server_submit copies the inputs and queues the request;
 * plom_server_step admits queued requests while their device bytes fit the budget
 * (FIFO), advances every admitted request by one interval in one batched launch round,
 * and retires finished ones; plom_server_status gives PLOM_SRV_QUEUED / RUNNING / DONE;
 * plom_server_result returns log-likelihood, n_all_fail, loglik_n and ess_n (and the
 * lineage of a PLOM_FLAG_LINEAGE request) of a done request and releases its device memory
 * (a finished request counts against the budget until then).  Requests whose work per
 * interval reaches the heavy threshold step in a lane of their own, concurrently with the
 * others, so that they do not hold back the small requests.  Requests take time-varying
 * covariates and drift maps like filters (checked and copied at submission).  Results equal
 * plom_smc_run on the same descriptor and inputs (FP64: bit for bit), whatever else is being
 * served. */
#define PLOM_SRV_QUEUED 0
#define PLOM_SRV_RUNNING 1
#define PLOM_SRV_DONE 2
typedef struct {
    int64_t hbm_budget_bytes;    /* 0: 80 % of the free device memory at creation */
    int32_t max_active;          /* 0: 4096 */
    int32_t max_inflight_steps;  /* 0: 4; per lane: a lane with that many steps queued waits */
    int64_t heavy_threshold;     /* J * size_proj from which a request runs in the heavy lane
                                    (its own stream, stepping independently of the small
                                    requests); 0: 4e6; negative: one lane for all */
    int32_t heavy_blocks;        /* while small requests are in flight, the heavy lane's
                                    propagate launches cover at most this many blocks each (tile
                                    ranges of one interval, one after another on its stream), so
                                    that the heavy request holds at most that many blocks' share of
                                    the CUs and the small requests' blocks find room at once;
                                    0: the device's CU count; negative: no limit */
    int32_t pad;
} plom_server_config;
typedef struct plom_server plom_server;
int plom_server_create(plom_ctx *ctx, const plom_server_config *cfg, plom_server **out);
int plom_server_submit(plom_server *srv, const plom_filter_desc *desc, const plom_filter_inputs *in, uint64_t *req_id);
int plom_server_step(plom_server *srv, int32_t *n_active, int32_t *n_queued, int32_t *n_done);
int plom_server_status(plom_server *srv, uint64_t req_id);
/* ids of the done requests whose results have not been taken, in completion order */
int plom_server_done(plom_server *srv, uint64_t *ids, int32_t cap, int32_t *n);
int plom_server_result(plom_server *srv, uint64_t req_id, plom_filter_outputs *out);
int plom_server_destroy(plom_server *srv);

int plom_sync(plom_ctx *ctx);

/* Timing of the last plom_filter_run / plom_batch_launch (eager) from HIP events.
 * total_ms always (context stream); with kernel timing on (plom_ctx_set_kernel_timing),
 * events also bracket the launches of every round: propagate_ms / scan_ms are the summed
 * durations of the propagate launches and of the weight / scan / advance launches,
 * n_intervals the number of launch rounds and n_propagate_launches the number of
 * propagate kernels they held (a batch launches one per (model, precision, variant)
 * group and lane in a round).  A batch on several lanes records each lane's events on
 * that lane's stream, so the durations are those of the launches as they run beside the
 * other lanes' launches; plom_last_timing_lanes gives the sums per lane.  No host
 * counterpart in the reference (its timing is the pmcmc/Tests.py wall clock). */
int plom_ctx_set_kernel_timing(plom_ctx *ctx, int on);
int plom_last_timing(plom_ctx *ctx, double *total_ms, double *propagate_ms, double *scan_ms, int32_t *n_intervals,
                     int32_t *n_propagate_launches);
/* per lane l < cap of the last timed run: summed propagate and weight / scan durations;
 * n_lanes = lanes that recorded events */
int plom_last_timing_lanes(plom_ctx *ctx, double *propagate_ms, double *scan_ms, int32_t cap, int32_t *n_lanes);
/* the device time the last timed run's propagate launches (and its weight / scan
 * launches) occupied: the length of the union of their event spans over all lanes, so
 * that launches running beside each other count once; -1 without timing */
int plom_last_timing_busy(plom_ctx *ctx, double *propagate_busy_ms, double *scan_busy_ms);

/* device self-test of the shared arithmetic: out[i*8+..] = exp, log, Phi, sin, cos, sqrt,
 * sqrt_mid (the Box-Muller radius square root), sqrt_rate (the demographic noise square
 * root) of in[i] */
int plom_selftest_math(plom_ctx *ctx, int n, const double *in, double *out);
/* plain device copy of nbytes (read nbytes + write nbytes), repeated: bandwidth reference */
int plom_selftest_copy(plom_ctx *ctx, int64_t nbytes, int reps, double *ms_per_copy);
/* device self-test of the streams: the first n normals of particle j at interval n0 */
int plom_selftest_normals(plom_ctx *ctx, uint32_t seed, uint32_t stream, uint32_t n0, uint32_t j,
                          int n, double *out);
/* the same Gaussians drawn by the FP32 stream (the same words, the transform in FP32),
 * widened to double */
int plom_selftest_normals32(plom_ctx *ctx, uint32_t seed, uint32_t stream, uint32_t n0, uint32_t j, int n, double *out);

#ifdef __cplusplus
}
#endif

#endif /* PLOM_SMC_H */
