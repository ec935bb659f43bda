/* Generated by iterated-inference-pipeline_amd/codegen from the sir model files. Do not edit. */
#ifndef PLOM_SIR_H
#define PLOM_SIR_H
#include "plom_model_common.h"

#define PLOM_SIR_N_CAC 1
#define PLOM_SIR_N_PAR_SV 2
#define PLOM_SIR_N_TS 1
#define PLOM_SIR_N_TS_INC_UNIQUE 1
#define PLOM_SIR_INC_OFFSET 2
#define PLOM_SIR_SIZE_PROJ 3
#define PLOM_SIR_PV_SIZE 5
#define PLOM_SIR_N_DEM 2
#define PLOM_SIR_N_ENV 0
#define PLOM_SIR_N_HAT 4
#define PLOM_SIR_N_PP 4
#define PLOM_SIR_FIX_OFF 4
#define PLOM_SIR_N_FIX 1
#define PLOM_SIR_N_DRIFT 0
#define PLOM_SIR_DRIFT_OFFSET 2

PLOM_HD void plom_sir_time_funcs_f64(double tr, const double *cvk, double *out)
{
    (void)tr; (void)cvk; (void)out;
    out[0] = (double)plom_recip_uniform(plom_cov_ld(cvk + 0));
}

PLOM_HD void plom_sir_interval_full_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const double P0 = PLOM_UNIFORM64(pv[0]);
    const double P1 = PLOM_UNIFORM64(pv[1]);
    X[2] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double RC0 = tft ? tft[k * 1 + 0] : plom_recip_uniform(C0);
        double s2 = 0.0;
        const double r0_0 = plom_div_rc((P0 * X[1]), C0, RC0);
        const double r1_0 = P1;
        double za0, zb0;
        plom_normal_pair64(ns, &za0, &zb0);
        const double d0_0 = sqdt * za0;
        const double d1_0 = sqdt * zb0;
        const double f0 = X[0] + ( - (r0_0*X[0]))*dtr + - plom_sqrt_rate((r0_0*X[0]))*d0_0;
        const double f1 = X[1] + ( - (r1_0*X[1]) + (r0_0*X[0]))*dtr + - plom_sqrt_rate((r1_0*X[1]))*d1_0+ plom_sqrt_rate((r0_0*X[0]))*d0_0;
        s2 += ( + (r0_0*X[0]))*dtr + + plom_sqrt_rate((r0_0*X[0]))*d0_0;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        const double f2 = X[2] + s2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        t += dt;
    }
}

PLOM_HD void plom_sir_interval_no_dem_sto_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const double P0 = PLOM_UNIFORM64(pv[0]);
    const double P1 = PLOM_UNIFORM64(pv[1]);
    X[2] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double RC0 = tft ? tft[k * 1 + 0] : plom_recip_uniform(C0);
        double s2 = 0.0;
        const double r0_0 = plom_div_rc((P0 * X[1]), C0, RC0);
        const double r1_0 = P1;
        const double f0 = X[0] + ( - (r0_0*X[0]))*dtr;
        const double f1 = X[1] + ( - (r1_0*X[1]) + (r0_0*X[0]))*dtr;
        s2 += ( + (r0_0*X[0]))*dtr;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        const double f2 = X[2] + s2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        t += dt;
    }
}

PLOM_HD void plom_sir_interval_no_env_sto_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const double P0 = PLOM_UNIFORM64(pv[0]);
    const double P1 = PLOM_UNIFORM64(pv[1]);
    X[2] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double RC0 = tft ? tft[k * 1 + 0] : plom_recip_uniform(C0);
        double s2 = 0.0;
        const double r0_0 = plom_div_rc((P0 * X[1]), C0, RC0);
        const double r1_0 = P1;
        double za0, zb0;
        plom_normal_pair64(ns, &za0, &zb0);
        const double d0_0 = sqdt * za0;
        const double d1_0 = sqdt * zb0;
        const double f0 = X[0] + ( - (r0_0*X[0]))*dtr + - plom_sqrt_rate((r0_0*X[0]))*d0_0;
        const double f1 = X[1] + ( - (r1_0*X[1]) + (r0_0*X[0]))*dtr + - plom_sqrt_rate((r1_0*X[1]))*d1_0+ plom_sqrt_rate((r0_0*X[0]))*d0_0;
        s2 += ( + (r0_0*X[0]))*dtr + + plom_sqrt_rate((r0_0*X[0]))*d0_0;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        const double f2 = X[2] + s2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        t += dt;
    }
}

PLOM_HD void plom_sir_interval_no_dem_sto_no_env_sto_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const double P0 = PLOM_UNIFORM64(pv[0]);
    const double P1 = PLOM_UNIFORM64(pv[1]);
    X[2] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double RC0 = tft ? tft[k * 1 + 0] : plom_recip_uniform(C0);
        double s2 = 0.0;
        const double r0_0 = plom_div_rc((P0 * X[1]), C0, RC0);
        const double r1_0 = P1;
        const double f0 = X[0] + ( - (r0_0*X[0]))*dtr;
        const double f1 = X[1] + ( - (r1_0*X[1]) + (r0_0*X[0]))*dtr;
        s2 += ( + (r0_0*X[0]))*dtr;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        const double f2 = X[2] + s2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        t += dt;
    }
}

PLOM_HD void plom_sir_proj2obs_f64(const double *X, const double *pvd, const double *cvd, double *obs)
{
    (void)pvd; (void)cvd; (void)X;
    obs[0] = (double)X[2];
}

PLOM_HD double plom_sir_loglik(const double *obs, const double *y, const double *pvd, const double *cvd, double t, double like_min, int *n_floor)
{
    double loglike = 0.0;
    int nf = 0;
    (void)t; (void)pvd; (void)cvd;
    if (!plom_isnan(y[0])) {
        const double x = obs[0];
        const double gsl_mu = (pvd[2] * x);
        const double gsl_sd = plom_sqrt((((pvd[2] * (1.0 - pvd[2])) * x) + plom_powi(((pvd[2] * pvd[3]) * x), 2)));
        double like;
        if (y[0] > 0.0) like = plom_cdf_gaussian_P(y[0] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[0] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[0] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    *n_floor = nf;
    return loglike;
}

PLOM_HD void plom_sir_obs_moments(const double *obs, const double *pvd, const double *cvd, double t, double *mean, double *var)
{
    (void)t; (void)pvd; (void)cvd;
    {
        const double x = obs[0];
        mean[0] = (pvd[2] * x);
        var[0] = (((pvd[2] * (1.0 - pvd[2])) * x) + plom_powi(((pvd[2] * pvd[3]) * x), 2));
    }
}

PLOM_HD void plom_sir_hat_values(const double *X, const double *pvd, const double *cvd, const plom_drift_view *dm, double t, double *out)
{
    (void)t; (void)pvd; (void)cvd; (void)dm;
    out[0] = X[0];
    out[1] = X[1];
    out[2] = cvd[0] - (X[0] + X[1]);
    double obs[1], var[1];
    plom_sir_proj2obs_f64(X, pvd, cvd, obs);
    plom_sir_obs_moments(obs, pvd, cvd, t, out + 3, var);
}

PLOM_HD void plom_sir_interval_psr_full_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const double P0 = PLOM_UNIFORM64(pv[0]);
    const double P1 = PLOM_UNIFORM64(pv[1]);
    X[2] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double RC0 = tft ? tft[k * 1 + 0] : plom_recip_uniform(C0);
        double n0_0 = 0.0;
        double n1_0 = 0.0;
        double stay0 = 0.0;
        double stay1 = 0.0;
        (void)n0_0; (void)n1_0;
        const double r0_0 = plom_div_rc((P0 * X[1]), C0, RC0);
        const double r1_0 = P1;
        {
            const double sum = r0_0 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r0_0 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[0]);
            n0_0 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay0 = (p0 == p0 && cnt > 0.0) ? cnt - n0_0 : 0.0;
        }
        {
            const double sum = r1_0 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_0 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[1]);
            n1_0 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay1 = (p0 == p0 && cnt > 0.0) ? cnt - n1_0 : 0.0;
        }
        X[0] = stay0;
        X[1] = stay1 + n0_0;
        X[2] = X[2] + (n0_0);
        t += dt;
    }
}

PLOM_HD void plom_sir_interval_psr_no_env_sto_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const double P0 = PLOM_UNIFORM64(pv[0]);
    const double P1 = PLOM_UNIFORM64(pv[1]);
    X[2] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double RC0 = tft ? tft[k * 1 + 0] : plom_recip_uniform(C0);
        double n0_0 = 0.0;
        double n1_0 = 0.0;
        double stay0 = 0.0;
        double stay1 = 0.0;
        (void)n0_0; (void)n1_0;
        const double r0_0 = plom_div_rc((P0 * X[1]), C0, RC0);
        const double r1_0 = P1;
        {
            const double sum = r0_0 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r0_0 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[0]);
            n0_0 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay0 = (p0 == p0 && cnt > 0.0) ? cnt - n0_0 : 0.0;
        }
        {
            const double sum = r1_0 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_0 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[1]);
            n1_0 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay1 = (p0 == p0 && cnt > 0.0) ? cnt - n1_0 : 0.0;
        }
        X[0] = stay0;
        X[1] = stay1 + n0_0;
        X[2] = X[2] + (n0_0);
        t += dt;
    }
}

PLOM_HD void plom_sir_ode_rhs(double t, const double *X, const double *pv, const double *lp, const double *cv, const plom_spline_set *sp, double *f)
{
    (void)pv; (void)lp; (void)cv; (void)sp;
    const double tr = t; (void)tr;
    const double P0 = PLOM_UNIFORM64(pv[0]);
    const double P1 = PLOM_UNIFORM64(pv[1]);
    const double C0 = sp ? plom_spline_at(sp, 0, t) : plom_cov_ld(cv + 0);
    const double r0_0 = ((P0 * X[1]) / C0);
    const double r1_0 = P1;
    f[0] =  - (r0_0*X[0]);
    f[1] =  - (r1_0*X[1]) + (r0_0*X[0]);
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r0_0*X[0]);
        f[2] = sum_inc;
    }
}

#define PLOM_ODE_FN plom_sir_ode_integrate
#define PLOM_ODE_RHS plom_sir_ode_rhs
#define PLOM_ODE_N 3
#define PLOM_ODE_QUAL PLOM_HD
#include "plom_ode_impl.h"

PLOM_HD int plom_sir_interval_ode_f64(double *X, const double *pv, const double *cv, const plom_spline_set *sp, double t0, double t1, double dt, double eps_abs, double eps_rel)
{
    X[2] = 0.0; /* reset_inc */
    return plom_sir_ode_integrate(X, pv, 0, cv, sp, t0, t1, dt, eps_abs, eps_rel);
}

PLOM_HD void plom_sir_interval_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, unsigned int noises_off, const double *cv, int cs, double eps_abs, double eps_rel, const plom_drift_view *dm, const plom_spline_set *sp)
{
    if (noises_off & PLOM_IMPL_ODE) {
        (void)plom_sir_interval_ode_f64(X, pv, cv, sp, t0, t1, dt, eps_abs, eps_rel);
        return;
    }
    if (noises_off & PLOM_IMPL_PSR) {
        if (noises_off & PLOM_NO_ENV_STO) plom_sir_interval_psr_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
        else plom_sir_interval_psr_full_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
        return;
    }
    const unsigned int v = noises_off & (PLOM_NO_DEM_STO | PLOM_NO_ENV_STO);
    if (v == 0u) plom_sir_interval_full_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
    else if (v == PLOM_NO_DEM_STO) plom_sir_interval_no_dem_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
    else if (v == PLOM_NO_ENV_STO) plom_sir_interval_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
    else plom_sir_interval_no_dem_sto_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
}

#if defined(__HIPCC__)
static __device__ inline __attribute__((always_inline)) void plom_sir_time_funcs_f32(float tr, const double *cvk, float *out)
{
    (void)tr; (void)cvk; (void)out;
    out[0] = (float)plom_recip_uniform(plom_cov_ld(cvk + 0));
}

static __device__ inline __attribute__((always_inline)) void plom_sir_interval_full_f32(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const float P0 = PLOM_UNIFORM32(pv[0]);
    const float P1 = PLOM_UNIFORM32(pv[1]);
    X[2] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        float s2 = 0.0f;
        const float r0_0 = plom_fdivf((P0 * X[1]), C0);
        const float r1_0 = P1;
        float za0, zb0;
        plom_normal_pair32(ns, &za0, &zb0);
        const float d0_0 = sqdt * za0;
        const float d1_0 = sqdt * zb0;
        const float f0 = X[0] + ( - (r0_0*X[0]))*dtr + - plom_fsqrtf((r0_0*X[0]))*d0_0;
        const float f1 = X[1] + ( - (r1_0*X[1]) + (r0_0*X[0]))*dtr + - plom_fsqrtf((r1_0*X[1]))*d1_0+ plom_fsqrtf((r0_0*X[0]))*d0_0;
        s2 += ( + (r0_0*X[0]))*dtr + + plom_fsqrtf((r0_0*X[0]))*d0_0;
        X[0] = plom_clipf(f0);
        X[1] = plom_clipf(f1);
        const float f2 = X[2] + s2;
        X[2] = plom_clipf(f2);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sir_interval_no_dem_sto_f32(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const float P0 = PLOM_UNIFORM32(pv[0]);
    const float P1 = PLOM_UNIFORM32(pv[1]);
    X[2] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        float s2 = 0.0f;
        const float r0_0 = plom_fdivf((P0 * X[1]), C0);
        const float r1_0 = P1;
        const float f0 = X[0] + ( - (r0_0*X[0]))*dtr;
        const float f1 = X[1] + ( - (r1_0*X[1]) + (r0_0*X[0]))*dtr;
        s2 += ( + (r0_0*X[0]))*dtr;
        X[0] = plom_clipf(f0);
        X[1] = plom_clipf(f1);
        const float f2 = X[2] + s2;
        X[2] = plom_clipf(f2);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sir_interval_no_env_sto_f32(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const float P0 = PLOM_UNIFORM32(pv[0]);
    const float P1 = PLOM_UNIFORM32(pv[1]);
    X[2] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        float s2 = 0.0f;
        const float r0_0 = plom_fdivf((P0 * X[1]), C0);
        const float r1_0 = P1;
        float za0, zb0;
        plom_normal_pair32(ns, &za0, &zb0);
        const float d0_0 = sqdt * za0;
        const float d1_0 = sqdt * zb0;
        const float f0 = X[0] + ( - (r0_0*X[0]))*dtr + - plom_fsqrtf((r0_0*X[0]))*d0_0;
        const float f1 = X[1] + ( - (r1_0*X[1]) + (r0_0*X[0]))*dtr + - plom_fsqrtf((r1_0*X[1]))*d1_0+ plom_fsqrtf((r0_0*X[0]))*d0_0;
        s2 += ( + (r0_0*X[0]))*dtr + + plom_fsqrtf((r0_0*X[0]))*d0_0;
        X[0] = plom_clipf(f0);
        X[1] = plom_clipf(f1);
        const float f2 = X[2] + s2;
        X[2] = plom_clipf(f2);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sir_interval_no_dem_sto_no_env_sto_f32(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const float P0 = PLOM_UNIFORM32(pv[0]);
    const float P1 = PLOM_UNIFORM32(pv[1]);
    X[2] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        float s2 = 0.0f;
        const float r0_0 = plom_fdivf((P0 * X[1]), C0);
        const float r1_0 = P1;
        const float f0 = X[0] + ( - (r0_0*X[0]))*dtr;
        const float f1 = X[1] + ( - (r1_0*X[1]) + (r0_0*X[0]))*dtr;
        s2 += ( + (r0_0*X[0]))*dtr;
        X[0] = plom_clipf(f0);
        X[1] = plom_clipf(f1);
        const float f2 = X[2] + s2;
        X[2] = plom_clipf(f2);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sir_proj2obs_f32(const float *X, const double *pvd, const double *cvd, double *obs)
{
    (void)pvd; (void)cvd; (void)X;
    obs[0] = (double)X[2];
}

static __device__ inline __attribute__((always_inline)) void plom_sir_interval_psr_full_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const double P0 = lp[0];
    const double P1 = lp[1];
    X[2] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double RC0 = tft ? tft[k * 1 + 0] : plom_recip_uniform(C0);
        double n0_0 = 0.0;
        double n1_0 = 0.0;
        double stay0 = 0.0;
        double stay1 = 0.0;
        (void)n0_0; (void)n1_0;
        const double r0_0 = plom_div_rc((P0 * X[1]), C0, RC0);
        const double r1_0 = P1;
        {
            const double sum = r0_0 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r0_0 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[0]);
            n0_0 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay0 = (p0 == p0 && cnt > 0.0) ? cnt - n0_0 : 0.0;
        }
        {
            const double sum = r1_0 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_0 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[1]);
            n1_0 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay1 = (p0 == p0 && cnt > 0.0) ? cnt - n1_0 : 0.0;
        }
        X[0] = stay0;
        X[1] = stay1 + n0_0;
        X[2] = X[2] + (n0_0);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sir_interval_psr_no_env_sto_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const double P0 = lp[0];
    const double P1 = lp[1];
    X[2] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double RC0 = tft ? tft[k * 1 + 0] : plom_recip_uniform(C0);
        double n0_0 = 0.0;
        double n1_0 = 0.0;
        double stay0 = 0.0;
        double stay1 = 0.0;
        (void)n0_0; (void)n1_0;
        const double r0_0 = plom_div_rc((P0 * X[1]), C0, RC0);
        const double r1_0 = P1;
        {
            const double sum = r0_0 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r0_0 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[0]);
            n0_0 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay0 = (p0 == p0 && cnt > 0.0) ? cnt - n0_0 : 0.0;
        }
        {
            const double sum = r1_0 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_0 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[1]);
            n1_0 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay1 = (p0 == p0 && cnt > 0.0) ? cnt - n1_0 : 0.0;
        }
        X[0] = stay0;
        X[1] = stay1 + n0_0;
        X[2] = X[2] + (n0_0);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sir_ode_rhs_pp(double t, const double *X, const double *pv, const double *lp, const double *cv, const plom_spline_set *sp, double *f)
{
    (void)pv; (void)lp; (void)cv; (void)sp;
    const double tr = t; (void)tr;
    const double P0 = lp[0];
    const double P1 = lp[1];
    const double C0 = sp ? plom_spline_at(sp, 0, t) : plom_cov_ld(cv + 0);
    const double r0_0 = ((P0 * X[1]) / C0);
    const double r1_0 = P1;
    f[0] =  - (r0_0*X[0]);
    f[1] =  - (r1_0*X[1]) + (r0_0*X[0]);
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r0_0*X[0]);
        f[2] = sum_inc;
    }
}

#define PLOM_ODE_FN plom_sir_ode_integrate_pp
#define PLOM_ODE_RHS plom_sir_ode_rhs_pp
#define PLOM_ODE_N 3
#define PLOM_ODE_QUAL static __device__ inline __attribute__((always_inline))
#include "plom_ode_impl.h"

static __device__ inline __attribute__((always_inline)) int plom_sir_interval_ode_pp_f64(double *X, const double *pv, const double *lp, const double *cv, const plom_spline_set *sp, double t0, double t1, double dt, double eps_abs, double eps_rel)
{
    X[2] = 0.0; /* reset_inc */
    return plom_sir_ode_integrate_pp(X, pv, lp, cv, sp, t0, t1, dt, eps_abs, eps_rel);
}

static __device__ inline __attribute__((always_inline)) void plom_sir_interval_full_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const double P0 = lp[0];
    const double P1 = lp[1];
    X[2] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double RC0 = tft ? tft[k * 1 + 0] : plom_recip_uniform(C0);
        double s2 = 0.0;
        const double r0_0 = plom_div_rc((P0 * X[1]), C0, RC0);
        const double r1_0 = P1;
        double za0, zb0;
        plom_normal_pair64(ns, &za0, &zb0);
        const double d0_0 = sqdt * za0;
        const double d1_0 = sqdt * zb0;
        const double f0 = X[0] + ( - (r0_0*X[0]))*dtr + - plom_sqrt_rate((r0_0*X[0]))*d0_0;
        const double f1 = X[1] + ( - (r1_0*X[1]) + (r0_0*X[0]))*dtr + - plom_sqrt_rate((r1_0*X[1]))*d1_0+ plom_sqrt_rate((r0_0*X[0]))*d0_0;
        s2 += ( + (r0_0*X[0]))*dtr + + plom_sqrt_rate((r0_0*X[0]))*d0_0;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        const double f2 = X[2] + s2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sir_interval_no_dem_sto_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const double P0 = lp[0];
    const double P1 = lp[1];
    X[2] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double RC0 = tft ? tft[k * 1 + 0] : plom_recip_uniform(C0);
        double s2 = 0.0;
        const double r0_0 = plom_div_rc((P0 * X[1]), C0, RC0);
        const double r1_0 = P1;
        const double f0 = X[0] + ( - (r0_0*X[0]))*dtr;
        const double f1 = X[1] + ( - (r1_0*X[1]) + (r0_0*X[0]))*dtr;
        s2 += ( + (r0_0*X[0]))*dtr;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        const double f2 = X[2] + s2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sir_interval_no_env_sto_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const double P0 = lp[0];
    const double P1 = lp[1];
    X[2] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double RC0 = tft ? tft[k * 1 + 0] : plom_recip_uniform(C0);
        double s2 = 0.0;
        const double r0_0 = plom_div_rc((P0 * X[1]), C0, RC0);
        const double r1_0 = P1;
        double za0, zb0;
        plom_normal_pair64(ns, &za0, &zb0);
        const double d0_0 = sqdt * za0;
        const double d1_0 = sqdt * zb0;
        const double f0 = X[0] + ( - (r0_0*X[0]))*dtr + - plom_sqrt_rate((r0_0*X[0]))*d0_0;
        const double f1 = X[1] + ( - (r1_0*X[1]) + (r0_0*X[0]))*dtr + - plom_sqrt_rate((r1_0*X[1]))*d1_0+ plom_sqrt_rate((r0_0*X[0]))*d0_0;
        s2 += ( + (r0_0*X[0]))*dtr + + plom_sqrt_rate((r0_0*X[0]))*d0_0;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        const double f2 = X[2] + s2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sir_interval_no_dem_sto_no_env_sto_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const double P0 = lp[0];
    const double P1 = lp[1];
    X[2] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double RC0 = tft ? tft[k * 1 + 0] : plom_recip_uniform(C0);
        double s2 = 0.0;
        const double r0_0 = plom_div_rc((P0 * X[1]), C0, RC0);
        const double r1_0 = P1;
        const double f0 = X[0] + ( - (r0_0*X[0]))*dtr;
        const double f1 = X[1] + ( - (r1_0*X[1]) + (r0_0*X[0]))*dtr;
        s2 += ( + (r0_0*X[0]))*dtr;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        const double f2 = X[2] + s2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sir_proj2obs_pp_f64(const double *X, const double *pvd, const double *lpd, const double *cvd, double *obs)
{
    (void)pvd; (void)lpd; (void)cvd; (void)X;
    obs[0] = (double)X[2];
}

static __device__ inline __attribute__((always_inline)) void plom_sir_interval_full_pp_f32(float *X, const float *pv, const float *lp, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const float P0 = lp[0];
    const float P1 = lp[1];
    X[2] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        float s2 = 0.0f;
        const float r0_0 = plom_fdivf((P0 * X[1]), C0);
        const float r1_0 = P1;
        float za0, zb0;
        plom_normal_pair32(ns, &za0, &zb0);
        const float d0_0 = sqdt * za0;
        const float d1_0 = sqdt * zb0;
        const float f0 = X[0] + ( - (r0_0*X[0]))*dtr + - plom_fsqrtf((r0_0*X[0]))*d0_0;
        const float f1 = X[1] + ( - (r1_0*X[1]) + (r0_0*X[0]))*dtr + - plom_fsqrtf((r1_0*X[1]))*d1_0+ plom_fsqrtf((r0_0*X[0]))*d0_0;
        s2 += ( + (r0_0*X[0]))*dtr + + plom_fsqrtf((r0_0*X[0]))*d0_0;
        X[0] = plom_clipf(f0);
        X[1] = plom_clipf(f1);
        const float f2 = X[2] + s2;
        X[2] = plom_clipf(f2);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sir_interval_no_dem_sto_pp_f32(float *X, const float *pv, const float *lp, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const float P0 = lp[0];
    const float P1 = lp[1];
    X[2] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        float s2 = 0.0f;
        const float r0_0 = plom_fdivf((P0 * X[1]), C0);
        const float r1_0 = P1;
        const float f0 = X[0] + ( - (r0_0*X[0]))*dtr;
        const float f1 = X[1] + ( - (r1_0*X[1]) + (r0_0*X[0]))*dtr;
        s2 += ( + (r0_0*X[0]))*dtr;
        X[0] = plom_clipf(f0);
        X[1] = plom_clipf(f1);
        const float f2 = X[2] + s2;
        X[2] = plom_clipf(f2);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sir_interval_no_env_sto_pp_f32(float *X, const float *pv, const float *lp, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const float P0 = lp[0];
    const float P1 = lp[1];
    X[2] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        float s2 = 0.0f;
        const float r0_0 = plom_fdivf((P0 * X[1]), C0);
        const float r1_0 = P1;
        float za0, zb0;
        plom_normal_pair32(ns, &za0, &zb0);
        const float d0_0 = sqdt * za0;
        const float d1_0 = sqdt * zb0;
        const float f0 = X[0] + ( - (r0_0*X[0]))*dtr + - plom_fsqrtf((r0_0*X[0]))*d0_0;
        const float f1 = X[1] + ( - (r1_0*X[1]) + (r0_0*X[0]))*dtr + - plom_fsqrtf((r1_0*X[1]))*d1_0+ plom_fsqrtf((r0_0*X[0]))*d0_0;
        s2 += ( + (r0_0*X[0]))*dtr + + plom_fsqrtf((r0_0*X[0]))*d0_0;
        X[0] = plom_clipf(f0);
        X[1] = plom_clipf(f1);
        const float f2 = X[2] + s2;
        X[2] = plom_clipf(f2);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sir_interval_no_dem_sto_no_env_sto_pp_f32(float *X, const float *pv, const float *lp, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const float P0 = lp[0];
    const float P1 = lp[1];
    X[2] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        float s2 = 0.0f;
        const float r0_0 = plom_fdivf((P0 * X[1]), C0);
        const float r1_0 = P1;
        const float f0 = X[0] + ( - (r0_0*X[0]))*dtr;
        const float f1 = X[1] + ( - (r1_0*X[1]) + (r0_0*X[0]))*dtr;
        s2 += ( + (r0_0*X[0]))*dtr;
        X[0] = plom_clipf(f0);
        X[1] = plom_clipf(f1);
        const float f2 = X[2] + s2;
        X[2] = plom_clipf(f2);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sir_proj2obs_pp_f32(const float *X, const double *pvd, const double *lpd, const double *cvd, double *obs)
{
    (void)pvd; (void)lpd; (void)cvd; (void)X;
    obs[0] = (double)X[2];
}

static __device__ inline __attribute__((always_inline)) double plom_sir_loglik_pp(const double *obs, const double *y, const double *pvd, const double *lpd, const double *cvd, double t, double like_min, int *n_floor)
{
    double loglike = 0.0;
    int nf = 0;
    (void)t; (void)pvd; (void)lpd; (void)cvd;
    if (!plom_isnan(y[0])) {
        const double x = obs[0];
        const double gsl_mu = (lpd[2] * x);
        const double gsl_sd = plom_sqrt((((lpd[2] * (1.0 - lpd[2])) * x) + plom_powi(((lpd[2] * lpd[3]) * x), 2)));
        double like;
        if (y[0] > 0.0) like = plom_cdf_gaussian_P(y[0] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[0] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[0] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    *n_floor = nf;
    return loglike;
}

struct plom_model_sir {
    static constexpr int S = 3, N_TS = 1, INC_OFFSET = 2, N_INC = 1, PV = 5, N_CAC = 1, N_HAT = 4, NPP = 4, FIX_OFF = 4, N_FIX = 1, N_DRIFT = 0;
    static constexpr int NTF = 1;  /* time-only functions and covariate reciprocals of a substep */
    template <int V> static __device__ inline __attribute__((always_inline)) void interval(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
    {
        if constexpr (V == 0) plom_sir_interval_full_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 1) plom_sir_interval_no_dem_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 2) plom_sir_interval_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 3) plom_sir_interval_no_dem_sto_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 4) plom_sir_interval_psr_full_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 5) plom_sir_interval_psr_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
    }
    template <int V> static __device__ inline __attribute__((always_inline)) void interval_pp(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
    {
        if constexpr (V == 0) plom_sir_interval_full_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 1) plom_sir_interval_no_dem_sto_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 2) plom_sir_interval_no_env_sto_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 3) plom_sir_interval_no_dem_sto_no_env_sto_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 4) plom_sir_interval_psr_full_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 5) plom_sir_interval_psr_no_env_sto_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
    }
    static __device__ inline __attribute__((always_inline)) void time_funcs(double tr, const double *cvk, double *out) { plom_sir_time_funcs_f64(tr, cvk, out); }
    static __device__ inline __attribute__((always_inline)) void proj2obs(const double *X, const double *pvd, const double *cvd, double *obs) { plom_sir_proj2obs_f64(X, pvd, cvd, obs); }
    static __device__ inline __attribute__((always_inline)) void proj2obs_pp(const double *X, const double *pvd, const double *lpd, const double *cvd, double *obs) { plom_sir_proj2obs_pp_f64(X, pvd, lpd, cvd, obs); }
    template <int V> static __device__ inline __attribute__((always_inline)) void interval(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
    {
        if constexpr (V == 0) plom_sir_interval_full_f32(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 1) plom_sir_interval_no_dem_sto_f32(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 2) plom_sir_interval_no_env_sto_f32(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 3) plom_sir_interval_no_dem_sto_no_env_sto_f32(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
    }
    template <int V> static __device__ inline __attribute__((always_inline)) void interval_pp(float *X, const float *pv, const float *lp, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
    {
        if constexpr (V == 0) plom_sir_interval_full_pp_f32(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 1) plom_sir_interval_no_dem_sto_pp_f32(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 2) plom_sir_interval_no_env_sto_pp_f32(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 3) plom_sir_interval_no_dem_sto_no_env_sto_pp_f32(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
    }
    static __device__ inline __attribute__((always_inline)) void time_funcs(float tr, const double *cvk, float *out) { plom_sir_time_funcs_f32(tr, cvk, out); }
    static __device__ inline __attribute__((always_inline)) void proj2obs(const float *X, const double *pvd, const double *cvd, double *obs) { plom_sir_proj2obs_f32(X, pvd, cvd, obs); }
    static __device__ inline __attribute__((always_inline)) void proj2obs_pp(const float *X, const double *pvd, const double *lpd, const double *cvd, double *obs) { plom_sir_proj2obs_pp_f32(X, pvd, lpd, cvd, obs); }
    static __device__ inline __attribute__((always_inline)) int interval_ode(double *X, const double *pv, const double *cv, const plom_spline_set *sp, double t0, double t1, double dt, double ea, double er) { return plom_sir_interval_ode_f64(X, pv, cv, sp, t0, t1, dt, ea, er); }
    static __device__ inline __attribute__((always_inline)) int interval_ode_pp(double *X, const double *pv, const double *lp, const double *cv, const plom_spline_set *sp, double t0, double t1, double dt, double ea, double er) { return plom_sir_interval_ode_pp_f64(X, pv, lp, cv, sp, t0, t1, dt, ea, er); }
    static __device__ inline __attribute__((always_inline)) double loglik(const double *obs, const double *y, const double *pvd, const double *cvd, double t, double like_min, int *n_floor) { return plom_sir_loglik(obs, y, pvd, cvd, t, like_min, n_floor); }
    static __device__ inline __attribute__((always_inline)) double loglik_pp(const double *obs, const double *y, const double *pvd, const double *lpd, const double *cvd, double t, double like_min, int *n_floor) { return plom_sir_loglik_pp(obs, y, pvd, lpd, cvd, t, like_min, n_floor); }
    static __device__ inline __attribute__((always_inline)) void hat_values(const double *X, const double *pvd, const double *cvd, const plom_drift_view *dm, double t, double *out) { plom_sir_hat_values(X, pvd, cvd, dm, t, out); }
    static __device__ inline __attribute__((always_inline)) void obs_moments(const double *obs, const double *pvd, const double *cvd, double t, double *mean, double *var) { plom_sir_obs_moments(obs, pvd, cvd, t, mean, var); }
};
#endif

#endif
