/*
 * plom_ode_impl.h -- defines one adaptive RKF45 interval function (method and contract in
 * plom_ode.h).  Included once per right-hand side, with PLOM_ODE_FN, PLOM_ODE_RHS,
 * PLOM_ODE_N and PLOM_ODE_QUAL defined; undefines them at the end.  The right-hand side is
 * called as PLOM_ODE_RHS(t, y, pv, lp, cv, sp, dydt): sp, when not null, gives the
 * covariates as splines of t, else cv holds their constant values.
 */
#include "plom_ode.h"

PLOM_ODE_QUAL int PLOM_ODE_FN(double *y, const double *pv, const double *lp, const double *cv, const plom_spline_set *sp,
                              double t0, double t1, double h, double eps_abs, double eps_rel)
{
    double k1[PLOM_ODE_N], k2[PLOM_ODE_N], k3[PLOM_ODE_N], k4[PLOM_ODE_N], k5[PLOM_ODE_N], k6[PLOM_ODE_N];
    double yt[PLOM_ODE_N], yn[PLOM_ODE_N];
    double t = t0;
    int steps = 0;
    while (t < t1) {
        /* derivative at the start of the step (used by every try from this state) */
        PLOM_ODE_RHS(t, y, pv, lp, cv, sp, k1);
        for (;;) {
            if (++steps > PLOM_ODE_MAX_STEPS) {
                for (int i = 0; i < PLOM_ODE_N; i++) y[i] = PLOM_NAN;
                return -1;
            }
            double hs = h;
            int last = 0;
            if (hs > t1 - t) {
                hs = t1 - t;
                last = 1;
            }
            for (int i = 0; i < PLOM_ODE_N; i++) yt[i] = y[i] + hs * (PLOM_RKF_A21 * k1[i]);
            PLOM_ODE_RHS(t + PLOM_RKF_C2 * hs, yt, pv, lp, cv, sp, k2);
            for (int i = 0; i < PLOM_ODE_N; i++) yt[i] = y[i] + hs * (PLOM_RKF_A31 * k1[i] + PLOM_RKF_A32 * k2[i]);
            PLOM_ODE_RHS(t + PLOM_RKF_C3 * hs, yt, pv, lp, cv, sp, k3);
            for (int i = 0; i < PLOM_ODE_N; i++)
                yt[i] = y[i] + hs * (PLOM_RKF_A41 * k1[i] + PLOM_RKF_A42 * k2[i] + PLOM_RKF_A43 * k3[i]);
            PLOM_ODE_RHS(t + PLOM_RKF_C4 * hs, yt, pv, lp, cv, sp, k4);
            for (int i = 0; i < PLOM_ODE_N; i++)
                yt[i] = y[i] + hs * (PLOM_RKF_A51 * k1[i] + PLOM_RKF_A52 * k2[i] + PLOM_RKF_A53 * k3[i] + PLOM_RKF_A54 * k4[i]);
            PLOM_ODE_RHS(t + hs, yt, pv, lp, cv, sp, k5);
            for (int i = 0; i < PLOM_ODE_N; i++)
                yt[i] = y[i] + hs * (PLOM_RKF_A61 * k1[i] + PLOM_RKF_A62 * k2[i] + PLOM_RKF_A63 * k3[i] +
                                     PLOM_RKF_A64 * k4[i] + PLOM_RKF_A65 * k5[i]);
            PLOM_ODE_RHS(t + PLOM_RKF_C6 * hs, yt, pv, lp, cv, sp, k6);
            /* fifth-order solution, error estimate and the largest error ratio */
            double rmax = PLOM_DBL_MIN;
            for (int i = 0; i < PLOM_ODE_N; i++) {
                yn[i] = y[i] + hs * (PLOM_RKF_B1 * k1[i] + PLOM_RKF_B3 * k3[i] + PLOM_RKF_B4 * k4[i] +
                                     PLOM_RKF_B5 * k5[i] + PLOM_RKF_B6 * k6[i]);
                const double err = hs * (PLOM_RKF_E1 * k1[i] + PLOM_RKF_E3 * k3[i] + PLOM_RKF_E4 * k4[i] +
                                         PLOM_RKF_E5 * k5[i] + PLOM_RKF_E6 * k6[i]);
                const double r = plom_fabs(err) / (eps_abs + eps_rel * plom_fabs(yn[i]));
                if (r > rmax) rmax = r;
            }
            const double tn = last ? t1 : t + hs;
            double hn = hs;
            if (rmax > 1.1) {
                double f = 0.9 / plom_pow(rmax, 1.0 / 5.0);
                if (f < 0.2) f = 0.2;
                hn = hs * f;
                if (plom_fabs(hn) < plom_fabs(hs) && tn + hn != tn) {
                    h = hn;          /* too large an error: the same step again, smaller */
                    continue;
                }
                hn = hs;
            } else if (rmax < 0.5) {
                double f = 0.9 / plom_pow(rmax, 1.0 / 6.0);
                if (f > 5.0) f = 5.0;
                if (f < 1.0) f = 1.0;
                hn = hs * f;
            }
            for (int i = 0; i < PLOM_ODE_N; i++) y[i] = yn[i];
            t = tn;
            h = hn;
            break;
        }
    }
    return 0;
}

#undef PLOM_ODE_FN
#undef PLOM_ODE_RHS
#undef PLOM_ODE_N
#undef PLOM_ODE_QUAL
