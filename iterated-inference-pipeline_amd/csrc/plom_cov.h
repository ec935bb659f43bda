/*
 * plom_cov.h -- time-varying covariates (the "metadata" of a model: population sizes,
 * birth rates, proportions under surveillance ...) as tables of values at the times where
 * the filter uses them.  Host code, shared by the library and the CPU oracle.
 *
 * The reference evaluates one interpolating spline per covariate and population (or
 * series) inside the generated code, at the start time t of every Euler substep and at
 * the observation time of the likelihood (gsl_spline_eval, model_builder/Ccoder.py:39-43;
 * set up in build_p_calc, C/core/build.c:1075-1211: gsl_interp_cspline above 3 knots,
 * gsl_interp_linear for 2 or 3).  Every particle of a filter sees the same times, so this
 * build evaluates each spline once per time on the host and the kernels read the table with
 * scalar loads.
 *
 *   plom_spline_prepare  natural cubic spline through (x_i, y_i): the second-derivative
 *                        coefficients c_i (c_0 = c_{n-1} = 0) from the tridiagonal system
 *                        h_{i-1} c_{i-1} + 2 (h_{i-1} + h_i) c_i + h_i c_{i+1}
 *                            = 3 ((y_{i+1} - y_i)/h_i - (y_i - y_{i-1})/h_{i-1}),
 *                        solved by forward elimination and back substitution (the system is
 *                        strictly diagonally dominant).
 *   plom_spline_eval     on the segment x_i <= t < x_{i+1} (the last segment for
 *                        t = x_{n-1}): y_i + dx (b_i + dx (c_i + dx d_i)) with
 *                        b_i = (y_{i+1} - y_i)/h_i - h_i (c_{i+1} + 2 c_i)/3 and
 *                        d_i = (c_{i+1} - c_i)/(3 h_i); linear interpolation for 2 or 3 knots;
 *                        the value itself for one knot (plom_spline_value of
 *                        plom_model_common.h, which the device evaluates too).
 *   plom_spline_pack     the knots and coefficients of every covariate in one array
 *                        (plom_spline_set), for the ode implementation.
 *   plom_cov_table       the table [rows][n_fix] of a filter (layout below).
 *
 * Table layout (n_fix = the model's fixed slots of pv, in pv order: every covariate per
 * population, then the covariates of the observation model per series):
 *   row n * kmax + k   (0 <= n < N, 0 <= k < K_n): the value at the start of substep k of
 *                       interval n, t = times[n] + dt + ... + dt (k terms, accumulated in
 *                       double as the generated code does);
 *   row N * kmax + r   (0 <= r <= N): the value at times[r] (the observation model and the
 *                       remainders of interval r - 1; r = 0 for the first row of hat).
 * kmax = max_n K_n.
 */
#ifndef PLOM_COV_H
#define PLOM_COV_H

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "plom_model_common.h"
#include "plom_smc.h"

/* c[n]: the coefficients of a natural cubic spline (n >= 4), or nothing (n < 4) */
static int plom_spline_prepare(int n, const double *x, const double *y, double *c)
{
    if (n < 4) return 0;
    const int m = n - 2; /* unknowns c_1 .. c_{n-2} */
    double *diag = (double *)malloc(sizeof(double) * (size_t)m);
    double *rhs = (double *)malloc(sizeof(double) * (size_t)m);
    if (!diag || !rhs) { free(diag); free(rhs); return -1; }
    for (int i = 1; i <= m; i++) {
        const double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
        diag[i - 1] = 2.0 * (h0 + h1);
        rhs[i - 1] = 3.0 * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
    }
    /* forward elimination: the sub- and super-diagonal of row i are h_{i-1} and h_i */
    for (int i = 1; i < m; i++) {
        const double h = x[i + 1] - x[i]; /* coupling of unknowns i and i + 1 (1-based) */
        const double f = h / diag[i - 1];
        diag[i] = diag[i] - f * h;
        rhs[i] = rhs[i] - f * rhs[i - 1];
    }
    c[0] = 0.0;
    c[n - 1] = 0.0;
    c[m] = rhs[m - 1] / diag[m - 1];
    for (int i = m - 1; i >= 1; i--) {
        const double h = x[i + 1] - x[i];
        c[i] = (rhs[i - 1] - h * c[i + 1]) / diag[i - 1];
    }
    free(diag);
    free(rhs);
    return 0;
}

static int plom_cov_check(const plom_covariate *cov, int n_fix, char *err, size_t errlen);

/* the value at t, x[0] <= t <= x[n-1] (checked by the caller) */
static double plom_spline_eval(int n, const double *x, const double *y, const double *c, double t)
{
    return plom_spline_value(n, x, y, c, t);
}

/* the splines of n_fix covariates packed for the device (plom_spline_set): n_out[n_fix],
 * off_out[n_fix], val_out[3 * total knots]; 0, or -1 on a bad series / no memory */
static int plom_spline_pack(const plom_covariate *cov, int n_fix, int32_t *n_out, int32_t *off_out, double *val_out,
                            char *err, size_t errlen)
{
    if (plom_cov_check(cov, n_fix, err, errlen)) return -1;
    size_t o = 0;
    for (int i = 0; i < n_fix; i++) {
        const int n = cov[i].n;
        n_out[i] = n;
        off_out[i] = (int32_t)o;
        memcpy(val_out + o, cov[i].x, sizeof(double) * (size_t)n);
        memcpy(val_out + o + n, cov[i].y, sizeof(double) * (size_t)n);
        for (int q = 0; q < n; q++) val_out[o + 2 * (size_t)n + q] = 0.0;
        if (plom_spline_prepare(n, cov[i].x, cov[i].y, val_out + o + 2 * (size_t)n)) {
            snprintf(err, errlen, "covariate %d: out of host memory", i);
            return -1;
        }
        o += 3 * (size_t)n;
    }
    return 0;
}

static size_t plom_spline_vals(const plom_covariate *cov, int n_fix)
{
    size_t v = 0;
    for (int i = 0; i < n_fix; i++) v += 3 * (size_t)cov[i].n;
    return v;
}

/* substep count of the longest interval */
static int plom_cov_kmax(int N, const double *times, double dt)
{
    int kmax = 0;
    for (int n = 0; n < N; n++) {
        const int K = plom_substeps(times[n], times[n + 1], dt);
        if (K > kmax) kmax = K;
    }
    return kmax;
}

static size_t plom_cov_rows(int N, int kmax) { return (size_t)N * (size_t)kmax + (size_t)N + 1; }

/* checks the series: n >= 1, finite values, strictly increasing knots */
static int plom_cov_check(const plom_covariate *cov, int n_fix, char *err, size_t errlen)
{
    for (int i = 0; i < n_fix; i++) {
        const plom_covariate *s = &cov[i];
        if (s->n < 1 || !s->x || !s->y) {
            snprintf(err, errlen, "covariate %d: needs at least one knot", i);
            return -1;
        }
        for (int q = 0; q < s->n; q++) {
            if (!(s->x[q] - s->x[q] == 0.0) || !(s->y[q] - s->y[q] == 0.0)) {
                snprintf(err, errlen, "covariate %d: knot %d is not finite", i, q);
                return -1;
            }
            if (q > 0 && !(s->x[q] > s->x[q - 1])) {
                snprintf(err, errlen, "covariate %d: knot times must increase strictly", i);
                return -1;
            }
        }
    }
    return 0;
}

/* out[plom_cov_rows(N, kmax) * n_fix]; 0, or -1 with a message when a time falls outside a
 * series' knots (the reference's spline evaluation fails there too) */
static int plom_cov_table(const plom_covariate *cov, int n_fix, int N, const double *times, double dt, int kmax,
                          double *out, char *err, size_t errlen)
{
    if (plom_cov_check(cov, n_fix, err, errlen)) return -1;
    const size_t rows = plom_cov_rows(N, kmax);
    for (int i = 0; i < n_fix; i++) {
        const plom_covariate *s = &cov[i];
        double *c = (double *)malloc(sizeof(double) * (size_t)(s->n + 1));
        if (!c || plom_spline_prepare(s->n, s->x, s->y, c)) {
            free(c);
            snprintf(err, errlen, "covariate %d: out of host memory", i);
            return -1;
        }
        const double lo = s->x[0], hi = s->x[s->n - 1];
        for (size_t r = 0; r < rows; r++) {
            double t;
            if (r >= (size_t)N * (size_t)kmax) {
                t = times[r - (size_t)N * (size_t)kmax];
            } else {
                const int n = (int)(r / (size_t)kmax), k = (int)(r % (size_t)kmax);
                const int K = plom_substeps(times[n], times[n + 1], dt);
                if (k >= K) { out[r * (size_t)n_fix + (size_t)i] = 0.0; continue; } /* padding */
                t = times[n];
                for (int q = 0; q < k; q++) t += dt;
            }
            if (s->n > 1 && !(t >= lo && t <= hi)) {
                free(c);
                snprintf(err, errlen, "covariate %d: time %.17g outside its knots [%.17g, %.17g]", i, t, lo, hi);
                return -1;
            }
            out[r * (size_t)n_fix + (size_t)i] = plom_spline_eval(s->n, s->x, s->y, c, t);
        }
        free(c);
    }
    return 0;
}

#endif
