/*
 * plom_cli.h -- host-side logic shared by the `smc` and `pmcmc` commands: the model
 * settings and theta files, parameter routers and transforms, the parameter vector and
 * initial state handed to the library, and the CSV / JSON writers whose formats the
 * reference's commands define.
 *
 * The parameter vector theta lives on the transformed scale, one value per (parameter,
 * group), parameters in the order par_sv, par_proc, par_obs (p_it_all,
 * build.c:400-475; offsets as plom_add_offset_iterator).
 */
#ifndef PLOM_CLI_H
#define PLOM_CLI_H

#include <stdio.h>

#include "plom_smc.h"
#include "pjson.h"

typedef enum { TR_ID, TR_LOG, TR_LOGIT, TR_LOGIT_AB } cli_tr;
typedef enum { SC_ID, SC_POW10, SC_POW10_NEG } cli_sc;
typedef enum { PRIOR_FLAT, PRIOR_NORMAL, PRIOR_PSEUDO_UNIFORM } cli_prior;

typedef struct {
    char name[128];
    int p;              /* number of populations (or series) */
    int n_gp;
    int *map;           /* element -> group */
    char **group_name;
    double *min, *max, *guess;
    double *sd;         /* sd_transf of each group (0 where absent) */
    cli_tr *tr;         /* transform of each group (set_f_trans, transform.c:332-410) */
    cli_prior *prior;   /* prior of each group (load.c:330-338) */
    cli_sc sc;
    double multiplier;
    int is_duration;
    int is_remainder;
    int follow;         /* index of the parameter it follows, -1 if none (build.c:1540-1589) */
    pj *par;            /* the theta parameter object (the followed one for a follower; NULL for the remainder) */
} cli_router;

typedef struct {
    pj *settings, *theta;
    const char *model_name;
    int model_id;
    plom_model_info mi;
    int N_CAC, N_TS, N_DATA, N;
    int n_sv, n_proc, n_obs, n_par;
    cli_router *R;
    int *offset;        /* theta offset of each parameter */
    int n_theta;        /* total number of (parameter, group) values */
    int ind_theta_remainder;
    pj *par_sv, *cac_id, *ts_id;
    int n_fixed;
    double *fixed;      /* [n_fixed][N_CAC] covariates at t = 0; k = 0 is the population size */
    plom_covariate *cov; /* [n_fix] the covariate series of the pv slots (plom_covariate) */
    double *cov0;       /* [n_fix] their values at t = 0 */
    double *cov_raw0;   /* [n_fix] the values at t = 0 before -g freezes the series */
    int has_cov;        /* some covariate changes with time: the filters take cov */
    double *pop_t;      /* [N + 1][N_CAC] population sizes at the observation times */
    int n_drift;        /* drifting parameters (settings drift, build_drift, build.c:609-646) */
    int *drift_par;     /* [n_drift] router index of each drifting parameter */
    plom_drift_map dmap; /* per parameter and population: group, back-transform, volatility slot */
    int drift_on;       /* the drift runs (set by the command: not with --no_diff or ode) */
    double *times;      /* [N + 1] */
    double *data;       /* [N][N_TS], NaN = missing */
    double dt;
    const char *remainder;
    int has_rem;        /* a remainder state is written (not POP_SIZE_EQ_SUM_SV) */
    int *noise_par;     /* [n_par] 1 for the white-noise sd parameters (ind_noise_sd) */
    int psr;            /* implementation psr: initial states rounded (prop2Xpop_size) */
} cli_problem;

extern int CLI_QUIET;
extern double CLI_FREEZE;   /* -g/--freeze_forcing: covariates constant from that time on (-1: off) */
extern int CLI_TRACE_17G;   /* trace values with 17 significant digits instead of %g */
void cli_log(const char *msg);

/* the implementation argument of the commands (C/smc/main.c:223-238): 'psr' (also the
 * default without one), 'sde' or 'ode'; anything else prints the help and exits with 1 */
#define CLI_SDE 0
#define CLI_PSR 1
#define CLI_ODE 2
int cli_implementation(int argc, char **argv, const char *help);
/* the implementation's flags (PLOM_FLAG_PSR / PLOM_FLAG_ODE); 'ode' also turns every noise
 * off; psr and ode refuse --fp32 */
void cli_apply_implementation(int impl, int fp32, unsigned int *noises_off, uint32_t *flags);
void cli_die(const char *msg);

double cli_f_tr(cli_tr k, double x, double a, double b);
double cli_f_inv(cli_tr k, double x, double a, double b);
double cli_f_inv_der(cli_tr k, double x, double a, double b);

/* Reads ./.settings.json and one line of theta from stdin and builds the routers
 * (is_bayesian: uniform priors use the logit_ab transform, transform.c:332-350). */
void cli_load(cli_problem *P, int is_bayesian, int nb_obs, double dt);

/* theta from the guesses: transform_theta and the theta remainder (load.c:365-373). */
void cli_theta_from_guess(const cli_problem *P, double *theta);

/* plom_check_IC_assign_theta_remainder (proposal.c:87-114): number of populations whose
 * initial proportions (without the remainder) exceed 1; assigns the remainder otherwise. */
int cli_theta_remainder(const cli_problem *P, double *theta);

/* apply_following_constraints (proposal.c:21-35) */
void cli_follow(const cli_problem *P, double *theta);

/* back_transform_x (transform.c:462-480) of group g of parameter i */
double cli_natural(const cli_problem *P, int i, int g, const double *theta);
/* prior densities on the natural scale: flat on [a, b], normal_prior, pseudo_unif_prior
 * (metropolis_hastings_prior.c:95-132) */
double cli_prior_pdf(cli_prior kind, double x, double a, double b);
/* Cholesky factor (lower, row-major) of the n x n matrix s A; 0 when positive definite */
int cli_cholesky(const double *A, int n, double s, double *L);
/* theta's "covariance" into var [n_theta][n_theta] (load_covariance, load.c:413-470); 0 without one */
int cli_load_covariance(const cli_problem *P, double *var);
/* the N(0, sd^2) density at u */
double cli_gauss_pdf(double u, double sd);
/* log prior of theta over the estimated parameters (sd_transf > 0); 1 on a numerical failure */
int cli_log_prior(const cli_problem *P, const double *theta, double like_min, double *out);

/* the parameter vector in the generated layout and the initial state */
void cli_pack(const cli_problem *P, const double *theta, double *pv, double *x0);

/* <path><name>_<id>.csv */
FILE *cli_fopen(const char *path, const char *name, int id);

/* trace: header_trace / print_trace (print.c:189-207, 396-440) */
void cli_trace_header(const cli_problem *P, FILE *f);
void cli_trace_row(const cli_problem *P, FILE *f, int index, const double *theta, double log_like);

/* X: header_X / print_p_X (print.c:88-129, 271-348) for a lineage of N + 1 rows printed
 * last row first, as sample_traj_and_print (print.c:715-784) */
void cli_x_header(const cli_problem *P, FILE *f);
void cli_x_rows(const cli_problem *P, FILE *f, int index, const double *lineage, const double *lineage_obs);
/* one X row (print_p_X, print.c:271-348): index, time of row r, states, remainders, observation
 * means ob[2 ts], drifting parameters, drawn observations ob[2 ts + 1] */
void cli_x_row(const cli_problem *P, FILE *f, int index, int r, const double *x, const double *ob);

/* plom_print_done (print.c:33-67): guesses replaced by f_inv(theta), then the theta JSON
 * on stdout (pipe, no trailing newline) or in <path>.done_<id>.json */
void cli_done(const cli_problem *P, const double *theta, const char *path, int id, int pipe);

#endif
