/*
 * plom_model_common.h -- helpers used by the generated model headers.
 *
 * PLOM_NO_* are the reference's noise switches (model_builder/C/core/plom.h:66).
 * The special functions follow model_builder/C/core/special_functions.c:45-74
 * (slowstep returns its argument inside the ramp; the reference returns an undeclared
 * `t` there, special_functions.c:74, which does not compile).
 * plom_sanitize_likelihood follows model_builder/C/core/likelihood.c:77-92 and counts
 * the floored series so that the filter can apply the all-series-floored failure rule.
 */
#ifndef PLOM_MODEL_COMMON_H
#define PLOM_MODEL_COMMON_H

#include "plom_rng.h"
#include "plom_psr.h"
#include "plom_mif.h"

/* The drifting parameters of a filter (diffusions, C/core/drift.c:27-42), per drifting
 * parameter i and population c (index i * n_cac + c): the group of the population (its
 * populations share one path), the back-transform of the drift state to the natural value
 * (as plom_mif_natural), and the pv slot of the natural volatility of group g in
 * vol_slot[i * n_cac + g] (-1: none); ng[i] is the number of groups.  A null view: no drift
 * (the parameters keep their pv values). */
typedef struct {
    const int32_t *group, *ng, *tr, *sc, *dur, *vol_slot;
    const double *min, *max, *mult;
} plom_drift_view;

/* PLOM_UNIFORM64/32: a value every lane of the wave holds (the filter's parameters: a
 * wave never spans two filters), moved to scalar registers on the device */
#if defined(__HIP_DEVICE_COMPILE__)
static __device__ inline __attribute__((always_inline)) double plom_uniform64(double x)
{
    uint64_t u = plom_d2u(x);
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
    return plom_u2d(((uint64_t)hi << 32) | lo);
}
static __device__ inline __attribute__((always_inline)) float plom_uniform32(float x)
{
    union { float f; uint32_t u; } v;
    v.f = x;
    v.u = __builtin_amdgcn_readfirstlane(v.u);
    return v.f;
}
#define PLOM_UNIFORM64(x) plom_uniform64(x)
#define PLOM_UNIFORM32(x) plom_uniform32(x)
#define PLOM_UNIFORM_INT(x) ((int)__builtin_amdgcn_readfirstlane((uint32_t)(x)))
#else
#define PLOM_UNIFORM64(x) (x)
#define PLOM_UNIFORM32(x) (x)
#define PLOM_UNIFORM_INT(x) (x)
#endif

/* one covariate value of the current substep (every lane reads the same address: a scalar
 * load through the constant address space on the device) */
#if defined(__HIP_DEVICE_COMPILE__)
static __device__ inline __attribute__((always_inline)) double plom_cov_ld(const double *p)
{
    return *(const __attribute__((address_space(4))) double *)p;
}
#elif defined(__HIPCC__)
static __host__ __device__ inline double plom_cov_ld(const double *p) { return *p; }
#else
static inline double plom_cov_ld(const double *p) { return *p; }
#endif

/* Time-varying covariates as splines, for the ode implementation, whose integrator asks for
 * them at any time (the reference's step_ode evaluates gsl_spline_eval at the integrator's
 * t, model_builder/Ccoder.py:39-43).  Covariate i (a fixed slot of pv, in pv order) has
 * n[i] knots; its knot times, values and natural-spline coefficients (plom_cov.h) are
 * val[off[i] + q], val[off[i] + n[i] + q] and val[off[i] + 2 n[i] + q]. */
typedef struct {
    const int32_t *n, *off;
    const double *val;
} plom_spline_set;

/* one spline at t: on the segment x_i <= t < x_{i+1} (the first or last segment outside the
 * knots) y_i + dx (b_i + dx (c_i + dx d_i)) with b_i = (y_{i+1} - y_i)/h_i - h_i (c_{i+1} +
 * 2 c_i)/3 and d_i = (c_{i+1} - c_i)/(3 h_i); linear interpolation for 2 or 3 knots; the value
 * itself for one knot */
PLOM_HD double plom_spline_value(int n, const double *x, const double *y, const double *c, double t)
{
    if (n == 1) return y[0];
    int lo = 0, hi = n - 1;
    while (hi > lo + 1) {
        const int mid = (lo + hi) / 2;
        if (x[mid] > t) hi = mid;
        else lo = mid;
    }
    const double h = x[lo + 1] - x[lo];
    const double dx = t - x[lo];
    const double dy = y[lo + 1] - y[lo];
    if (n < 4) return y[lo] + dx * (dy / h);
    const double b = dy / h - h * (c[lo + 1] + 2.0 * c[lo]) / 3.0;
    const double d = (c[lo + 1] - c[lo]) / (3.0 * h);
    return y[lo] + dx * (b + dx * (c[lo] + dx * d));
}

PLOM_HD double plom_spline_at(const plom_spline_set *sp, int i, double t)
{
    const int n = sp->n[i];
    const double *b = sp->val + sp->off[i];
    return plom_spline_value(n, b, b + n, b + 2 * n, t);
}

/* number of Euler substeps of an interval: the iterations of
 * t = t0; while (t < t1) t += dt;  (prediction_util.c:275-287, t accumulated in double) */
PLOM_HD int plom_substeps(double t0, double t1, double dt)
{
    int K = 0;
    double t = t0;
    while (t < t1) {
        t += dt;
        K++;
    }
    return K;
}

#define PLOM_NO_DEM_STO 1u
#define PLOM_NO_ENV_STO 2u
/* the host dispatchers' selector of the psr implementation (Euler-multinomial step) */
#define PLOM_IMPL_PSR 0x100u
/* the host dispatchers' selector of the ode implementation (adaptive RKF45) */
#define PLOM_IMPL_ODE 0x200u
#define PLOM_NO_DRIFT 4u

PLOM_HD double plom_heaviside(double x) { return (x < 0.0) ? 0.0 : 1.0; }
PLOM_HD double plom_ramp(double x) { return (x >= 0.0) ? x : 0.0; }
PLOM_HD double plom_slowstep(double x, double d) { return (x >= 0.0) ? ((x >= d) ? d : x) : 0.0; }

/* School-term forcing, terms_forcing(amplitude) of a rate (C/core/special_functions.c:21-47).
 * rec is the population's school record in pv: n_terms, prop_school (the proportion of the
 * year with school in: 1 - sum(end - begin), C/core/build.c:860-866), then the holiday
 * periods (begin, end) in years.  With y = t/365 - floor(t/365) the fraction of the year at
 * time t (days), the forcing is amplitude / (1 + amplitude (1 - prop_school)) when
 * begin <= y <= end for some period, and 1 / (1 + amplitude (1 - prop_school)) otherwise. */
#ifndef PLOM_MAX_TERMS
#define PLOM_MAX_TERMS 8
#define PLOM_SCHOOL_REC (2 + 2 * PLOM_MAX_TERMS)
#endif
PLOM_HD double plom_terms_forcing(double amp, double t, const double *rec)
{
    const double ty = t / 365.0;
    const double y = ty - __builtin_floor(ty);
    const int n = (int)rec[0];
    int hol = 0;
    for (int k = 0; k < n && k < PLOM_MAX_TERMS; k++) hol |= (y >= rec[2 + 2 * k]) & (y <= rec[3 + 2 * k]);
    return (hol ? amp : 1.0) / (1.0 + amp * (1.0 - rec[1]));
}

PLOM_HD double plom_sanitize_likelihood(double like, double like_min, int *n_floor)
{
    if (plom_isnan(like) || like == PLOM_INF || like == -PLOM_INF || like < 0.0) {
        (*n_floor)++;
        return like_min;
    }
    if (like <= like_min) {
        (*n_floor)++;
        return like_min;
    }
    return like;
}

#if defined(__HIPCC__)
#define PLOM_DF static __device__ inline __attribute__((always_inline))
PLOM_DF float plom_sinf(float x) { return sinf(x); }
PLOM_DF float plom_cosf(float x) { return cosf(x); }
PLOM_DF float plom_expf(float x) { return expf(x); }
PLOM_DF float plom_logf(float x) { return logf(x); }
PLOM_DF float plom_sqrtf(float x) { return __builtin_sqrtf(x); }
PLOM_DF float plom_fabsf(float x) { return fabsf(x); }
PLOM_DF float plom_powif(float x, int n)
{
    int neg = n < 0;
    unsigned int m = neg ? (unsigned int)(-n) : (unsigned int)n;
    float r = 1.0f, b = x;
    while (m) {
        if (m & 1u) r = r * b;
        m >>= 1;
        if (m) b = b * b;
    }
    return neg ? 1.0f / r : r;
}
PLOM_DF float plom_powf(float x, float y) { return powf(x, y); }
PLOM_DF float plom_heavisidef(float x) { return (x < 0.0f) ? 0.0f : 1.0f; }
PLOM_DF float plom_rampf(float x) { return (x >= 0.0f) ? x : 0.0f; }
PLOM_DF float plom_slowstepf(float x, float d) { return (x >= 0.0f) ? ((x >= d) ? d : x) : 0.0f; }
PLOM_DF float plom_terms_forcingf(float amp, float t, const float *rec)
{
    const float ty = t / 365.0f;
    const float y = ty - __builtin_floorf(ty);
    const int n = (int)rec[0];
    int hol = 0;
    for (int k = 0; k < n && k < PLOM_MAX_TERMS; k++) hol |= (y >= rec[2 + 2 * k]) & (y <= rec[3 + 2 * k]);
    return (hol ? amp : 1.0f) / (1.0f + amp * (1.0f - rec[1]));
}
/* FP32 propagation is its own Monte Carlo estimate (checked against FP64 statistically),
 * so its noise square roots, divisions and clipping use the single-instruction hardware
 * forms (v_sqrt_f32 and v_rcp_f32 within 1 ulp, v_max_f32) instead of the correctly
 * rounded expansions */
PLOM_DF float plom_fsqrtf(float x) { return __builtin_amdgcn_sqrtf(x); }
PLOM_DF float plom_fdivf(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
PLOM_DF float plom_clipf(float f) { return __builtin_fmaxf(f, 0.0f); }
#endif

#endif /* PLOM_MODEL_COMMON_H */
