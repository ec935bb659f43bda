/*
 * plom_ode.h -- the deterministic implementation "ode": adaptive Runge-Kutta-Fehlberg 4(5)
 * integration of a model's rate equations over one observation interval, shared by the
 * GPU kernels and the CPU oracle (compiled with -ffp-contract=off on both sides).
 *
 * The reference integrates the generated step_ode (C/templates/core_template.c:215-339,
 * derivative form without clipping) with GSL: gsl_odeiv2_step_rkf45 under
 * gsl_odeiv2_control_y_new(eps_abs, eps_rel), driven by gsl_odeiv2_evolve_apply from
 * t = t0 with an initial step h = dt until t1 (f_prediction_ode,
 * C/core/prediction_util.c:176-193; defaults eps_abs 1e-6, eps_rel 1e-3, C/core/plom.h:90-91).
 * GSL is not part of the reference checkout (no pinned version); this header restates the
 * published methods:
 *
 *   step     Fehlberg's embedded 4(5) pair (E. Fehlberg, NASA TR R-315, 1969): six stages,
 *            the fifth-order combination advances the solution and the difference of the
 *            fifth- and fourth-order combinations estimates the error (local extrapolation).
 *   control  the GSL manual's standard control with a_y = 1, a_dydt = 0: desired error
 *            D_i = eps_abs + eps_rel |y_i| (of the new solution), observed E_i = |yerr_i|,
 *            r = max_i E_i / D_i (at least the smallest normal double).  r > 1.1: the step
 *            size becomes h 0.9 r^(-1/5), at least h / 5, and the step is taken again from
 *            the same state, unless the smaller step would not move t (then it is kept);
 *            r < 0.5: the next step size is h 0.9 r^(-1/6), at most 5 h; otherwise h stays.
 *            The exponents use the order 5 of the advancing solution.
 *   evolve   a step never passes t1: a step size beyond t1 - t is cut to t1 - t and lands on
 *            t1 exactly; the step size carries over from step to step within the interval.
 *
 * A step count beyond PLOM_ODE_MAX_STEPS in one interval (tried steps, accepted or not) ends
 * the integration with a NaN state, which the likelihood floors; a GPU lane always leaves
 * the loop.
 *
 * Usage: a generated model header defines
 *   PLOM_ODE_FN   the interval function to define,
 *   PLOM_ODE_RHS  its right-hand side f = dy/dt:
 *                 void rhs(double t, const double *y, const double *pv, const double *lp,
 *                          const double *cv, double *f),
 *   PLOM_ODE_N    the dimension, PLOM_ODE_QUAL the function qualifiers,
 * then includes plom_ode_impl.h, which defines
 *   int PLOM_ODE_FN(double *y, const double *pv, const double *lp, const double *cv,
 *                   double t0, double t1, double h, double eps_abs, double eps_rel);
 * (0, or -1 when the step bound was reached).
 */
#ifndef PLOM_ODE_H
#define PLOM_ODE_H

#include "plom_math.h"

#define PLOM_ODE_MAX_STEPS 100000
#define PLOM_ODE_EPS_ABS 1e-6
#define PLOM_ODE_EPS_REL 1e-3
#define PLOM_DBL_MIN 2.2250738585072014e-308

/* Fehlberg's coefficients: nodes c_i, stage weights a_ij, the fifth-order weights b_i and
 * the error weights e_i = b_i - b*_i (b* the fourth-order weights) */
#define PLOM_RKF_C2 (1.0 / 4.0)
#define PLOM_RKF_C3 (3.0 / 8.0)
#define PLOM_RKF_C4 (12.0 / 13.0)
#define PLOM_RKF_C6 (1.0 / 2.0)
#define PLOM_RKF_A21 (1.0 / 4.0)
#define PLOM_RKF_A31 (3.0 / 32.0)
#define PLOM_RKF_A32 (9.0 / 32.0)
#define PLOM_RKF_A41 (1932.0 / 2197.0)
#define PLOM_RKF_A42 (-7200.0 / 2197.0)
#define PLOM_RKF_A43 (7296.0 / 2197.0)
#define PLOM_RKF_A51 (439.0 / 216.0)
#define PLOM_RKF_A52 (-8.0)
#define PLOM_RKF_A53 (3680.0 / 513.0)
#define PLOM_RKF_A54 (-845.0 / 4104.0)
#define PLOM_RKF_A61 (-8.0 / 27.0)
#define PLOM_RKF_A62 (2.0)
#define PLOM_RKF_A63 (-3544.0 / 2565.0)
#define PLOM_RKF_A64 (1859.0 / 4104.0)
#define PLOM_RKF_A65 (-11.0 / 40.0)
#define PLOM_RKF_B1 (16.0 / 135.0)
#define PLOM_RKF_B3 (6656.0 / 12825.0)
#define PLOM_RKF_B4 (28561.0 / 56430.0)
#define PLOM_RKF_B5 (-9.0 / 50.0)
#define PLOM_RKF_B6 (2.0 / 55.0)
#define PLOM_RKF_E1 (1.0 / 360.0)
#define PLOM_RKF_E3 (-128.0 / 4275.0)
#define PLOM_RKF_E4 (-2197.0 / 75240.0)
#define PLOM_RKF_E5 (1.0 / 50.0)
#define PLOM_RKF_E6 (2.0 / 55.0)

#endif
