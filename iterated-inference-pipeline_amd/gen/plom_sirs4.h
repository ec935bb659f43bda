/* Generated by iterated-inference-pipeline_amd/codegen from the sirs4 model files. Do not edit. */
#ifndef PLOM_SIRS4_H
#define PLOM_SIRS4_H
#include "plom_model_common.h"

#define PLOM_SIRS4_N_CAC 4
#define PLOM_SIRS4_N_PAR_SV 2
#define PLOM_SIRS4_N_TS 4
#define PLOM_SIRS4_N_TS_INC_UNIQUE 4
#define PLOM_SIRS4_INC_OFFSET 8
#define PLOM_SIRS4_SIZE_PROJ 12
#define PLOM_SIRS4_PV_SIZE 48
#define PLOM_SIRS4_N_DEM 3
#define PLOM_SIRS4_N_ENV 1

PLOM_HD void plom_sirs4_interval_full_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv;
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    double t = t0;
    while (t < t1) {
        const double tr = (double)t; (void)tr;
        const double r0_0 = (((pv[0] * (1.0 + (pv[4] * plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0))))) * (pv[32] * X[4] + pv[33] * X[5] + pv[34] * X[6] + pv[35] * X[7])) / pv[28]);
        const double r1_0 = pv[8];
        const double r2_0 = (pv[12] * ((pv[28] - X[0]) - X[4]));
        const double d0_0 = sqdt * plom_normal64(ns);
        const double d1_0 = sqdt * plom_normal64(ns);
        const double d2_0 = sqdt * plom_normal64(ns);
        const double w0_0 = sqdt * plom_normal64(ns);
        const double r0_1 = (((pv[1] * (1.0 + (pv[5] * plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0))))) * (pv[36] * X[4] + pv[37] * X[5] + pv[38] * X[6] + pv[39] * X[7])) / pv[29]);
        const double r1_1 = pv[9];
        const double r2_1 = (pv[13] * ((pv[29] - X[1]) - X[5]));
        const double d0_1 = sqdt * plom_normal64(ns);
        const double d1_1 = sqdt * plom_normal64(ns);
        const double d2_1 = sqdt * plom_normal64(ns);
        const double w0_1 = sqdt * plom_normal64(ns);
        const double r0_2 = (((pv[2] * (1.0 + (pv[6] * plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0))))) * (pv[40] * X[4] + pv[41] * X[5] + pv[42] * X[6] + pv[43] * X[7])) / pv[30]);
        const double r1_2 = pv[10];
        const double r2_2 = (pv[14] * ((pv[30] - X[2]) - X[6]));
        const double d0_2 = sqdt * plom_normal64(ns);
        const double d1_2 = sqdt * plom_normal64(ns);
        const double d2_2 = sqdt * plom_normal64(ns);
        const double w0_2 = sqdt * plom_normal64(ns);
        const double r0_3 = (((pv[3] * (1.0 + (pv[7] * plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0))))) * (pv[44] * X[4] + pv[45] * X[5] + pv[46] * X[6] + pv[47] * X[7])) / pv[31]);
        const double r1_3 = pv[11];
        const double r2_3 = (pv[15] * ((pv[31] - X[3]) - X[7]));
        const double d0_3 = sqdt * plom_normal64(ns);
        const double d1_3 = sqdt * plom_normal64(ns);
        const double d2_3 = sqdt * plom_normal64(ns);
        const double w0_3 = sqdt * plom_normal64(ns);
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - plom_sqrt((r0_0*X[0]))*d0_0+ plom_sqrt((r2_0))*d2_0 + - (r0_0*X[0])*pv[16]*w0_0;
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - plom_sqrt((r0_1*X[1]))*d0_1+ plom_sqrt((r2_1))*d2_1 + - (r0_1*X[1])*pv[17]*w0_1;
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - plom_sqrt((r0_2*X[2]))*d0_2+ plom_sqrt((r2_2))*d2_2 + - (r0_2*X[2])*pv[18]*w0_2;
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - plom_sqrt((r0_3*X[3]))*d0_3+ plom_sqrt((r2_3))*d2_3 + - (r0_3*X[3])*pv[19]*w0_3;
        const double f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + - plom_sqrt((r1_0*X[4]))*d1_0+ plom_sqrt((r0_0*X[0]))*d0_0 + + (r0_0*X[0])*pv[16]*w0_0;
        const double f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + - plom_sqrt((r1_1*X[5]))*d1_1+ plom_sqrt((r0_1*X[1]))*d0_1 + + (r0_1*X[1])*pv[17]*w0_1;
        const double f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + - plom_sqrt((r1_2*X[6]))*d1_2+ plom_sqrt((r0_2*X[2]))*d0_2 + + (r0_2*X[2])*pv[18]*w0_2;
        const double f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + - plom_sqrt((r1_3*X[7]))*d1_3+ plom_sqrt((r0_3*X[3]))*d0_3 + + (r0_3*X[3])*pv[19]*w0_3;
        double s8 = 0.0;
        s8 += ( + (r0_0*X[0]))*dtr + + plom_sqrt((r0_0*X[0]))*d0_0 + + (r0_0*X[0])*pv[16]*w0_0;
        const double f8 = X[8] + s8;
        double s9 = 0.0;
        s9 += ( + (r0_1*X[1]))*dtr + + plom_sqrt((r0_1*X[1]))*d0_1 + + (r0_1*X[1])*pv[17]*w0_1;
        const double f9 = X[9] + s9;
        double s10 = 0.0;
        s10 += ( + (r0_2*X[2]))*dtr + + plom_sqrt((r0_2*X[2]))*d0_2 + + (r0_2*X[2])*pv[18]*w0_2;
        const double f10 = X[10] + s10;
        double s11 = 0.0;
        s11 += ( + (r0_3*X[3]))*dtr + + plom_sqrt((r0_3*X[3]))*d0_3 + + (r0_3*X[3])*pv[19]*w0_3;
        const double f11 = X[11] + s11;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        t += dt;
    }
}

PLOM_HD void plom_sirs4_interval_no_dem_sto_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv;
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    double t = t0;
    while (t < t1) {
        const double tr = (double)t; (void)tr;
        const double r0_0 = (((pv[0] * (1.0 + (pv[4] * plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0))))) * (pv[32] * X[4] + pv[33] * X[5] + pv[34] * X[6] + pv[35] * X[7])) / pv[28]);
        const double r1_0 = pv[8];
        const double r2_0 = (pv[12] * ((pv[28] - X[0]) - X[4]));
        const double w0_0 = sqdt * plom_normal64(ns);
        const double r0_1 = (((pv[1] * (1.0 + (pv[5] * plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0))))) * (pv[36] * X[4] + pv[37] * X[5] + pv[38] * X[6] + pv[39] * X[7])) / pv[29]);
        const double r1_1 = pv[9];
        const double r2_1 = (pv[13] * ((pv[29] - X[1]) - X[5]));
        const double w0_1 = sqdt * plom_normal64(ns);
        const double r0_2 = (((pv[2] * (1.0 + (pv[6] * plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0))))) * (pv[40] * X[4] + pv[41] * X[5] + pv[42] * X[6] + pv[43] * X[7])) / pv[30]);
        const double r1_2 = pv[10];
        const double r2_2 = (pv[14] * ((pv[30] - X[2]) - X[6]));
        const double w0_2 = sqdt * plom_normal64(ns);
        const double r0_3 = (((pv[3] * (1.0 + (pv[7] * plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0))))) * (pv[44] * X[4] + pv[45] * X[5] + pv[46] * X[6] + pv[47] * X[7])) / pv[31]);
        const double r1_3 = pv[11];
        const double r2_3 = (pv[15] * ((pv[31] - X[3]) - X[7]));
        const double w0_3 = sqdt * plom_normal64(ns);
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - (r0_0*X[0])*pv[16]*w0_0;
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - (r0_1*X[1])*pv[17]*w0_1;
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - (r0_2*X[2])*pv[18]*w0_2;
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - (r0_3*X[3])*pv[19]*w0_3;
        const double f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + + (r0_0*X[0])*pv[16]*w0_0;
        const double f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + + (r0_1*X[1])*pv[17]*w0_1;
        const double f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + + (r0_2*X[2])*pv[18]*w0_2;
        const double f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + + (r0_3*X[3])*pv[19]*w0_3;
        double s8 = 0.0;
        s8 += ( + (r0_0*X[0]))*dtr + + (r0_0*X[0])*pv[16]*w0_0;
        const double f8 = X[8] + s8;
        double s9 = 0.0;
        s9 += ( + (r0_1*X[1]))*dtr + + (r0_1*X[1])*pv[17]*w0_1;
        const double f9 = X[9] + s9;
        double s10 = 0.0;
        s10 += ( + (r0_2*X[2]))*dtr + + (r0_2*X[2])*pv[18]*w0_2;
        const double f10 = X[10] + s10;
        double s11 = 0.0;
        s11 += ( + (r0_3*X[3]))*dtr + + (r0_3*X[3])*pv[19]*w0_3;
        const double f11 = X[11] + s11;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        t += dt;
    }
}

PLOM_HD void plom_sirs4_interval_no_env_sto_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv;
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    double t = t0;
    while (t < t1) {
        const double tr = (double)t; (void)tr;
        const double r0_0 = (((pv[0] * (1.0 + (pv[4] * plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0))))) * (pv[32] * X[4] + pv[33] * X[5] + pv[34] * X[6] + pv[35] * X[7])) / pv[28]);
        const double r1_0 = pv[8];
        const double r2_0 = (pv[12] * ((pv[28] - X[0]) - X[4]));
        const double d0_0 = sqdt * plom_normal64(ns);
        const double d1_0 = sqdt * plom_normal64(ns);
        const double d2_0 = sqdt * plom_normal64(ns);
        const double r0_1 = (((pv[1] * (1.0 + (pv[5] * plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0))))) * (pv[36] * X[4] + pv[37] * X[5] + pv[38] * X[6] + pv[39] * X[7])) / pv[29]);
        const double r1_1 = pv[9];
        const double r2_1 = (pv[13] * ((pv[29] - X[1]) - X[5]));
        const double d0_1 = sqdt * plom_normal64(ns);
        const double d1_1 = sqdt * plom_normal64(ns);
        const double d2_1 = sqdt * plom_normal64(ns);
        const double r0_2 = (((pv[2] * (1.0 + (pv[6] * plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0))))) * (pv[40] * X[4] + pv[41] * X[5] + pv[42] * X[6] + pv[43] * X[7])) / pv[30]);
        const double r1_2 = pv[10];
        const double r2_2 = (pv[14] * ((pv[30] - X[2]) - X[6]));
        const double d0_2 = sqdt * plom_normal64(ns);
        const double d1_2 = sqdt * plom_normal64(ns);
        const double d2_2 = sqdt * plom_normal64(ns);
        const double r0_3 = (((pv[3] * (1.0 + (pv[7] * plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0))))) * (pv[44] * X[4] + pv[45] * X[5] + pv[46] * X[6] + pv[47] * X[7])) / pv[31]);
        const double r1_3 = pv[11];
        const double r2_3 = (pv[15] * ((pv[31] - X[3]) - X[7]));
        const double d0_3 = sqdt * plom_normal64(ns);
        const double d1_3 = sqdt * plom_normal64(ns);
        const double d2_3 = sqdt * plom_normal64(ns);
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - plom_sqrt((r0_0*X[0]))*d0_0+ plom_sqrt((r2_0))*d2_0;
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - plom_sqrt((r0_1*X[1]))*d0_1+ plom_sqrt((r2_1))*d2_1;
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - plom_sqrt((r0_2*X[2]))*d0_2+ plom_sqrt((r2_2))*d2_2;
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - plom_sqrt((r0_3*X[3]))*d0_3+ plom_sqrt((r2_3))*d2_3;
        const double f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + - plom_sqrt((r1_0*X[4]))*d1_0+ plom_sqrt((r0_0*X[0]))*d0_0;
        const double f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + - plom_sqrt((r1_1*X[5]))*d1_1+ plom_sqrt((r0_1*X[1]))*d0_1;
        const double f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + - plom_sqrt((r1_2*X[6]))*d1_2+ plom_sqrt((r0_2*X[2]))*d0_2;
        const double f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + - plom_sqrt((r1_3*X[7]))*d1_3+ plom_sqrt((r0_3*X[3]))*d0_3;
        double s8 = 0.0;
        s8 += ( + (r0_0*X[0]))*dtr + + plom_sqrt((r0_0*X[0]))*d0_0;
        const double f8 = X[8] + s8;
        double s9 = 0.0;
        s9 += ( + (r0_1*X[1]))*dtr + + plom_sqrt((r0_1*X[1]))*d0_1;
        const double f9 = X[9] + s9;
        double s10 = 0.0;
        s10 += ( + (r0_2*X[2]))*dtr + + plom_sqrt((r0_2*X[2]))*d0_2;
        const double f10 = X[10] + s10;
        double s11 = 0.0;
        s11 += ( + (r0_3*X[3]))*dtr + + plom_sqrt((r0_3*X[3]))*d0_3;
        const double f11 = X[11] + s11;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        t += dt;
    }
}

PLOM_HD void plom_sirs4_interval_no_dem_sto_no_env_sto_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv;
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    double t = t0;
    while (t < t1) {
        const double tr = (double)t; (void)tr;
        const double r0_0 = (((pv[0] * (1.0 + (pv[4] * plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0))))) * (pv[32] * X[4] + pv[33] * X[5] + pv[34] * X[6] + pv[35] * X[7])) / pv[28]);
        const double r1_0 = pv[8];
        const double r2_0 = (pv[12] * ((pv[28] - X[0]) - X[4]));
        const double r0_1 = (((pv[1] * (1.0 + (pv[5] * plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0))))) * (pv[36] * X[4] + pv[37] * X[5] + pv[38] * X[6] + pv[39] * X[7])) / pv[29]);
        const double r1_1 = pv[9];
        const double r2_1 = (pv[13] * ((pv[29] - X[1]) - X[5]));
        const double r0_2 = (((pv[2] * (1.0 + (pv[6] * plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0))))) * (pv[40] * X[4] + pv[41] * X[5] + pv[42] * X[6] + pv[43] * X[7])) / pv[30]);
        const double r1_2 = pv[10];
        const double r2_2 = (pv[14] * ((pv[30] - X[2]) - X[6]));
        const double r0_3 = (((pv[3] * (1.0 + (pv[7] * plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0))))) * (pv[44] * X[4] + pv[45] * X[5] + pv[46] * X[6] + pv[47] * X[7])) / pv[31]);
        const double r1_3 = pv[11];
        const double r2_3 = (pv[15] * ((pv[31] - X[3]) - X[7]));
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr;
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr;
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr;
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr;
        const double f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr;
        const double f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr;
        const double f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr;
        const double f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr;
        double s8 = 0.0;
        s8 += ( + (r0_0*X[0]))*dtr;
        const double f8 = X[8] + s8;
        double s9 = 0.0;
        s9 += ( + (r0_1*X[1]))*dtr;
        const double f9 = X[9] + s9;
        double s10 = 0.0;
        s10 += ( + (r0_2*X[2]))*dtr;
        const double f10 = X[10] + s10;
        double s11 = 0.0;
        s11 += ( + (r0_3*X[3]))*dtr;
        const double f11 = X[11] + s11;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        t += dt;
    }
}

PLOM_HD void plom_sirs4_proj2obs_f64(const double *X, const double *pvd, double *obs)
{
    (void)pvd; (void)X;
    obs[0] = (double)X[8];
    obs[1] = (double)X[9];
    obs[2] = (double)X[10];
    obs[3] = (double)X[11];
}

PLOM_HD double plom_sirs4_loglik(const double *obs, const double *y, const double *pvd, double t, double like_min, int *n_floor)
{
    double loglike = 0.0;
    int nf = 0;
    (void)t; (void)pvd;
    if (!plom_isnan(y[0])) {
        const double x = obs[0];
        const double gsl_mu = (pvd[20] * x);
        const double gsl_sd = plom_sqrt((((pvd[20] * (1.0 - pvd[20])) * x) + plom_powi(((pvd[20] * pvd[24]) * x), 2)));
        double like;
        if (y[0] > 0.0) like = plom_cdf_gaussian_P(y[0] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[0] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[0] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[1])) {
        const double x = obs[1];
        const double gsl_mu = (pvd[21] * x);
        const double gsl_sd = plom_sqrt((((pvd[21] * (1.0 - pvd[21])) * x) + plom_powi(((pvd[21] * pvd[25]) * x), 2)));
        double like;
        if (y[1] > 0.0) like = plom_cdf_gaussian_P(y[1] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[1] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[1] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[2])) {
        const double x = obs[2];
        const double gsl_mu = (pvd[22] * x);
        const double gsl_sd = plom_sqrt((((pvd[22] * (1.0 - pvd[22])) * x) + plom_powi(((pvd[22] * pvd[26]) * x), 2)));
        double like;
        if (y[2] > 0.0) like = plom_cdf_gaussian_P(y[2] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[2] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[2] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[3])) {
        const double x = obs[3];
        const double gsl_mu = (pvd[23] * x);
        const double gsl_sd = plom_sqrt((((pvd[23] * (1.0 - pvd[23])) * x) + plom_powi(((pvd[23] * pvd[27]) * x), 2)));
        double like;
        if (y[3] > 0.0) like = plom_cdf_gaussian_P(y[3] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[3] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[3] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    *n_floor = nf;
    return loglike;
}

PLOM_HD void plom_sirs4_interval_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, unsigned int noises_off)
{
    const unsigned int v = noises_off & (PLOM_NO_DEM_STO | PLOM_NO_ENV_STO);
    if (v == 0u) plom_sirs4_interval_full_f64(X, pv, ns, t0, t1, dt, sqdt);
    else if (v == PLOM_NO_DEM_STO) plom_sirs4_interval_no_dem_sto_f64(X, pv, ns, t0, t1, dt, sqdt);
    else if (v == PLOM_NO_ENV_STO) plom_sirs4_interval_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt);
    else plom_sirs4_interval_no_dem_sto_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt);
}

#if defined(__HIPCC__)
static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_full_f32(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv;
    X[8] = 0.0f; /* reset_inc */
    X[9] = 0.0f; /* reset_inc */
    X[10] = 0.0f; /* reset_inc */
    X[11] = 0.0f; /* reset_inc */
    double t = t0;
    while (t < t1) {
        const float tr = (float)t; (void)tr;
        const float r0_0 = (((pv[0] * (1.0f + (pv[4] * plom_sinf((((2.0f * 3.141592653589793f) * tr) / 365.0f))))) * (pv[32] * X[4] + pv[33] * X[5] + pv[34] * X[6] + pv[35] * X[7])) / pv[28]);
        const float r1_0 = pv[8];
        const float r2_0 = (pv[12] * ((pv[28] - X[0]) - X[4]));
        const float d0_0 = sqdt * plom_normal32(ns);
        const float d1_0 = sqdt * plom_normal32(ns);
        const float d2_0 = sqdt * plom_normal32(ns);
        const float w0_0 = sqdt * plom_normal32(ns);
        const float r0_1 = (((pv[1] * (1.0f + (pv[5] * plom_sinf((((2.0f * 3.141592653589793f) * tr) / 365.0f))))) * (pv[36] * X[4] + pv[37] * X[5] + pv[38] * X[6] + pv[39] * X[7])) / pv[29]);
        const float r1_1 = pv[9];
        const float r2_1 = (pv[13] * ((pv[29] - X[1]) - X[5]));
        const float d0_1 = sqdt * plom_normal32(ns);
        const float d1_1 = sqdt * plom_normal32(ns);
        const float d2_1 = sqdt * plom_normal32(ns);
        const float w0_1 = sqdt * plom_normal32(ns);
        const float r0_2 = (((pv[2] * (1.0f + (pv[6] * plom_sinf((((2.0f * 3.141592653589793f) * tr) / 365.0f))))) * (pv[40] * X[4] + pv[41] * X[5] + pv[42] * X[6] + pv[43] * X[7])) / pv[30]);
        const float r1_2 = pv[10];
        const float r2_2 = (pv[14] * ((pv[30] - X[2]) - X[6]));
        const float d0_2 = sqdt * plom_normal32(ns);
        const float d1_2 = sqdt * plom_normal32(ns);
        const float d2_2 = sqdt * plom_normal32(ns);
        const float w0_2 = sqdt * plom_normal32(ns);
        const float r0_3 = (((pv[3] * (1.0f + (pv[7] * plom_sinf((((2.0f * 3.141592653589793f) * tr) / 365.0f))))) * (pv[44] * X[4] + pv[45] * X[5] + pv[46] * X[6] + pv[47] * X[7])) / pv[31]);
        const float r1_3 = pv[11];
        const float r2_3 = (pv[15] * ((pv[31] - X[3]) - X[7]));
        const float d0_3 = sqdt * plom_normal32(ns);
        const float d1_3 = sqdt * plom_normal32(ns);
        const float d2_3 = sqdt * plom_normal32(ns);
        const float w0_3 = sqdt * plom_normal32(ns);
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - __builtin_sqrtf((r0_0*X[0]))*d0_0+ __builtin_sqrtf((r2_0))*d2_0 + - (r0_0*X[0])*pv[16]*w0_0;
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - __builtin_sqrtf((r0_1*X[1]))*d0_1+ __builtin_sqrtf((r2_1))*d2_1 + - (r0_1*X[1])*pv[17]*w0_1;
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - __builtin_sqrtf((r0_2*X[2]))*d0_2+ __builtin_sqrtf((r2_2))*d2_2 + - (r0_2*X[2])*pv[18]*w0_2;
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - __builtin_sqrtf((r0_3*X[3]))*d0_3+ __builtin_sqrtf((r2_3))*d2_3 + - (r0_3*X[3])*pv[19]*w0_3;
        const float f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + - __builtin_sqrtf((r1_0*X[4]))*d1_0+ __builtin_sqrtf((r0_0*X[0]))*d0_0 + + (r0_0*X[0])*pv[16]*w0_0;
        const float f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + - __builtin_sqrtf((r1_1*X[5]))*d1_1+ __builtin_sqrtf((r0_1*X[1]))*d0_1 + + (r0_1*X[1])*pv[17]*w0_1;
        const float f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + - __builtin_sqrtf((r1_2*X[6]))*d1_2+ __builtin_sqrtf((r0_2*X[2]))*d0_2 + + (r0_2*X[2])*pv[18]*w0_2;
        const float f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + - __builtin_sqrtf((r1_3*X[7]))*d1_3+ __builtin_sqrtf((r0_3*X[3]))*d0_3 + + (r0_3*X[3])*pv[19]*w0_3;
        float s8 = 0.0f;
        s8 += ( + (r0_0*X[0]))*dtr + + __builtin_sqrtf((r0_0*X[0]))*d0_0 + + (r0_0*X[0])*pv[16]*w0_0;
        const float f8 = X[8] + s8;
        float s9 = 0.0f;
        s9 += ( + (r0_1*X[1]))*dtr + + __builtin_sqrtf((r0_1*X[1]))*d0_1 + + (r0_1*X[1])*pv[17]*w0_1;
        const float f9 = X[9] + s9;
        float s10 = 0.0f;
        s10 += ( + (r0_2*X[2]))*dtr + + __builtin_sqrtf((r0_2*X[2]))*d0_2 + + (r0_2*X[2])*pv[18]*w0_2;
        const float f10 = X[10] + s10;
        float s11 = 0.0f;
        s11 += ( + (r0_3*X[3]))*dtr + + __builtin_sqrtf((r0_3*X[3]))*d0_3 + + (r0_3*X[3])*pv[19]*w0_3;
        const float f11 = X[11] + s11;
        X[0] = (f0 < 0.0f) ? 0.0f : f0;
        X[1] = (f1 < 0.0f) ? 0.0f : f1;
        X[2] = (f2 < 0.0f) ? 0.0f : f2;
        X[3] = (f3 < 0.0f) ? 0.0f : f3;
        X[4] = (f4 < 0.0f) ? 0.0f : f4;
        X[5] = (f5 < 0.0f) ? 0.0f : f5;
        X[6] = (f6 < 0.0f) ? 0.0f : f6;
        X[7] = (f7 < 0.0f) ? 0.0f : f7;
        X[8] = (f8 < 0.0f) ? 0.0f : f8;
        X[9] = (f9 < 0.0f) ? 0.0f : f9;
        X[10] = (f10 < 0.0f) ? 0.0f : f10;
        X[11] = (f11 < 0.0f) ? 0.0f : f11;
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_no_dem_sto_f32(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv;
    X[8] = 0.0f; /* reset_inc */
    X[9] = 0.0f; /* reset_inc */
    X[10] = 0.0f; /* reset_inc */
    X[11] = 0.0f; /* reset_inc */
    double t = t0;
    while (t < t1) {
        const float tr = (float)t; (void)tr;
        const float r0_0 = (((pv[0] * (1.0f + (pv[4] * plom_sinf((((2.0f * 3.141592653589793f) * tr) / 365.0f))))) * (pv[32] * X[4] + pv[33] * X[5] + pv[34] * X[6] + pv[35] * X[7])) / pv[28]);
        const float r1_0 = pv[8];
        const float r2_0 = (pv[12] * ((pv[28] - X[0]) - X[4]));
        const float w0_0 = sqdt * plom_normal32(ns);
        const float r0_1 = (((pv[1] * (1.0f + (pv[5] * plom_sinf((((2.0f * 3.141592653589793f) * tr) / 365.0f))))) * (pv[36] * X[4] + pv[37] * X[5] + pv[38] * X[6] + pv[39] * X[7])) / pv[29]);
        const float r1_1 = pv[9];
        const float r2_1 = (pv[13] * ((pv[29] - X[1]) - X[5]));
        const float w0_1 = sqdt * plom_normal32(ns);
        const float r0_2 = (((pv[2] * (1.0f + (pv[6] * plom_sinf((((2.0f * 3.141592653589793f) * tr) / 365.0f))))) * (pv[40] * X[4] + pv[41] * X[5] + pv[42] * X[6] + pv[43] * X[7])) / pv[30]);
        const float r1_2 = pv[10];
        const float r2_2 = (pv[14] * ((pv[30] - X[2]) - X[6]));
        const float w0_2 = sqdt * plom_normal32(ns);
        const float r0_3 = (((pv[3] * (1.0f + (pv[7] * plom_sinf((((2.0f * 3.141592653589793f) * tr) / 365.0f))))) * (pv[44] * X[4] + pv[45] * X[5] + pv[46] * X[6] + pv[47] * X[7])) / pv[31]);
        const float r1_3 = pv[11];
        const float r2_3 = (pv[15] * ((pv[31] - X[3]) - X[7]));
        const float w0_3 = sqdt * plom_normal32(ns);
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - (r0_0*X[0])*pv[16]*w0_0;
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - (r0_1*X[1])*pv[17]*w0_1;
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - (r0_2*X[2])*pv[18]*w0_2;
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - (r0_3*X[3])*pv[19]*w0_3;
        const float f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + + (r0_0*X[0])*pv[16]*w0_0;
        const float f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + + (r0_1*X[1])*pv[17]*w0_1;
        const float f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + + (r0_2*X[2])*pv[18]*w0_2;
        const float f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + + (r0_3*X[3])*pv[19]*w0_3;
        float s8 = 0.0f;
        s8 += ( + (r0_0*X[0]))*dtr + + (r0_0*X[0])*pv[16]*w0_0;
        const float f8 = X[8] + s8;
        float s9 = 0.0f;
        s9 += ( + (r0_1*X[1]))*dtr + + (r0_1*X[1])*pv[17]*w0_1;
        const float f9 = X[9] + s9;
        float s10 = 0.0f;
        s10 += ( + (r0_2*X[2]))*dtr + + (r0_2*X[2])*pv[18]*w0_2;
        const float f10 = X[10] + s10;
        float s11 = 0.0f;
        s11 += ( + (r0_3*X[3]))*dtr + + (r0_3*X[3])*pv[19]*w0_3;
        const float f11 = X[11] + s11;
        X[0] = (f0 < 0.0f) ? 0.0f : f0;
        X[1] = (f1 < 0.0f) ? 0.0f : f1;
        X[2] = (f2 < 0.0f) ? 0.0f : f2;
        X[3] = (f3 < 0.0f) ? 0.0f : f3;
        X[4] = (f4 < 0.0f) ? 0.0f : f4;
        X[5] = (f5 < 0.0f) ? 0.0f : f5;
        X[6] = (f6 < 0.0f) ? 0.0f : f6;
        X[7] = (f7 < 0.0f) ? 0.0f : f7;
        X[8] = (f8 < 0.0f) ? 0.0f : f8;
        X[9] = (f9 < 0.0f) ? 0.0f : f9;
        X[10] = (f10 < 0.0f) ? 0.0f : f10;
        X[11] = (f11 < 0.0f) ? 0.0f : f11;
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_no_env_sto_f32(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv;
    X[8] = 0.0f; /* reset_inc */
    X[9] = 0.0f; /* reset_inc */
    X[10] = 0.0f; /* reset_inc */
    X[11] = 0.0f; /* reset_inc */
    double t = t0;
    while (t < t1) {
        const float tr = (float)t; (void)tr;
        const float r0_0 = (((pv[0] * (1.0f + (pv[4] * plom_sinf((((2.0f * 3.141592653589793f) * tr) / 365.0f))))) * (pv[32] * X[4] + pv[33] * X[5] + pv[34] * X[6] + pv[35] * X[7])) / pv[28]);
        const float r1_0 = pv[8];
        const float r2_0 = (pv[12] * ((pv[28] - X[0]) - X[4]));
        const float d0_0 = sqdt * plom_normal32(ns);
        const float d1_0 = sqdt * plom_normal32(ns);
        const float d2_0 = sqdt * plom_normal32(ns);
        const float r0_1 = (((pv[1] * (1.0f + (pv[5] * plom_sinf((((2.0f * 3.141592653589793f) * tr) / 365.0f))))) * (pv[36] * X[4] + pv[37] * X[5] + pv[38] * X[6] + pv[39] * X[7])) / pv[29]);
        const float r1_1 = pv[9];
        const float r2_1 = (pv[13] * ((pv[29] - X[1]) - X[5]));
        const float d0_1 = sqdt * plom_normal32(ns);
        const float d1_1 = sqdt * plom_normal32(ns);
        const float d2_1 = sqdt * plom_normal32(ns);
        const float r0_2 = (((pv[2] * (1.0f + (pv[6] * plom_sinf((((2.0f * 3.141592653589793f) * tr) / 365.0f))))) * (pv[40] * X[4] + pv[41] * X[5] + pv[42] * X[6] + pv[43] * X[7])) / pv[30]);
        const float r1_2 = pv[10];
        const float r2_2 = (pv[14] * ((pv[30] - X[2]) - X[6]));
        const float d0_2 = sqdt * plom_normal32(ns);
        const float d1_2 = sqdt * plom_normal32(ns);
        const float d2_2 = sqdt * plom_normal32(ns);
        const float r0_3 = (((pv[3] * (1.0f + (pv[7] * plom_sinf((((2.0f * 3.141592653589793f) * tr) / 365.0f))))) * (pv[44] * X[4] + pv[45] * X[5] + pv[46] * X[6] + pv[47] * X[7])) / pv[31]);
        const float r1_3 = pv[11];
        const float r2_3 = (pv[15] * ((pv[31] - X[3]) - X[7]));
        const float d0_3 = sqdt * plom_normal32(ns);
        const float d1_3 = sqdt * plom_normal32(ns);
        const float d2_3 = sqdt * plom_normal32(ns);
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - __builtin_sqrtf((r0_0*X[0]))*d0_0+ __builtin_sqrtf((r2_0))*d2_0;
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - __builtin_sqrtf((r0_1*X[1]))*d0_1+ __builtin_sqrtf((r2_1))*d2_1;
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - __builtin_sqrtf((r0_2*X[2]))*d0_2+ __builtin_sqrtf((r2_2))*d2_2;
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - __builtin_sqrtf((r0_3*X[3]))*d0_3+ __builtin_sqrtf((r2_3))*d2_3;
        const float f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + - __builtin_sqrtf((r1_0*X[4]))*d1_0+ __builtin_sqrtf((r0_0*X[0]))*d0_0;
        const float f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + - __builtin_sqrtf((r1_1*X[5]))*d1_1+ __builtin_sqrtf((r0_1*X[1]))*d0_1;
        const float f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + - __builtin_sqrtf((r1_2*X[6]))*d1_2+ __builtin_sqrtf((r0_2*X[2]))*d0_2;
        const float f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + - __builtin_sqrtf((r1_3*X[7]))*d1_3+ __builtin_sqrtf((r0_3*X[3]))*d0_3;
        float s8 = 0.0f;
        s8 += ( + (r0_0*X[0]))*dtr + + __builtin_sqrtf((r0_0*X[0]))*d0_0;
        const float f8 = X[8] + s8;
        float s9 = 0.0f;
        s9 += ( + (r0_1*X[1]))*dtr + + __builtin_sqrtf((r0_1*X[1]))*d0_1;
        const float f9 = X[9] + s9;
        float s10 = 0.0f;
        s10 += ( + (r0_2*X[2]))*dtr + + __builtin_sqrtf((r0_2*X[2]))*d0_2;
        const float f10 = X[10] + s10;
        float s11 = 0.0f;
        s11 += ( + (r0_3*X[3]))*dtr + + __builtin_sqrtf((r0_3*X[3]))*d0_3;
        const float f11 = X[11] + s11;
        X[0] = (f0 < 0.0f) ? 0.0f : f0;
        X[1] = (f1 < 0.0f) ? 0.0f : f1;
        X[2] = (f2 < 0.0f) ? 0.0f : f2;
        X[3] = (f3 < 0.0f) ? 0.0f : f3;
        X[4] = (f4 < 0.0f) ? 0.0f : f4;
        X[5] = (f5 < 0.0f) ? 0.0f : f5;
        X[6] = (f6 < 0.0f) ? 0.0f : f6;
        X[7] = (f7 < 0.0f) ? 0.0f : f7;
        X[8] = (f8 < 0.0f) ? 0.0f : f8;
        X[9] = (f9 < 0.0f) ? 0.0f : f9;
        X[10] = (f10 < 0.0f) ? 0.0f : f10;
        X[11] = (f11 < 0.0f) ? 0.0f : f11;
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_no_dem_sto_no_env_sto_f32(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv;
    X[8] = 0.0f; /* reset_inc */
    X[9] = 0.0f; /* reset_inc */
    X[10] = 0.0f; /* reset_inc */
    X[11] = 0.0f; /* reset_inc */
    double t = t0;
    while (t < t1) {
        const float tr = (float)t; (void)tr;
        const float r0_0 = (((pv[0] * (1.0f + (pv[4] * plom_sinf((((2.0f * 3.141592653589793f) * tr) / 365.0f))))) * (pv[32] * X[4] + pv[33] * X[5] + pv[34] * X[6] + pv[35] * X[7])) / pv[28]);
        const float r1_0 = pv[8];
        const float r2_0 = (pv[12] * ((pv[28] - X[0]) - X[4]));
        const float r0_1 = (((pv[1] * (1.0f + (pv[5] * plom_sinf((((2.0f * 3.141592653589793f) * tr) / 365.0f))))) * (pv[36] * X[4] + pv[37] * X[5] + pv[38] * X[6] + pv[39] * X[7])) / pv[29]);
        const float r1_1 = pv[9];
        const float r2_1 = (pv[13] * ((pv[29] - X[1]) - X[5]));
        const float r0_2 = (((pv[2] * (1.0f + (pv[6] * plom_sinf((((2.0f * 3.141592653589793f) * tr) / 365.0f))))) * (pv[40] * X[4] + pv[41] * X[5] + pv[42] * X[6] + pv[43] * X[7])) / pv[30]);
        const float r1_2 = pv[10];
        const float r2_2 = (pv[14] * ((pv[30] - X[2]) - X[6]));
        const float r0_3 = (((pv[3] * (1.0f + (pv[7] * plom_sinf((((2.0f * 3.141592653589793f) * tr) / 365.0f))))) * (pv[44] * X[4] + pv[45] * X[5] + pv[46] * X[6] + pv[47] * X[7])) / pv[31]);
        const float r1_3 = pv[11];
        const float r2_3 = (pv[15] * ((pv[31] - X[3]) - X[7]));
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr;
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr;
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr;
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr;
        const float f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr;
        const float f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr;
        const float f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr;
        const float f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr;
        float s8 = 0.0f;
        s8 += ( + (r0_0*X[0]))*dtr;
        const float f8 = X[8] + s8;
        float s9 = 0.0f;
        s9 += ( + (r0_1*X[1]))*dtr;
        const float f9 = X[9] + s9;
        float s10 = 0.0f;
        s10 += ( + (r0_2*X[2]))*dtr;
        const float f10 = X[10] + s10;
        float s11 = 0.0f;
        s11 += ( + (r0_3*X[3]))*dtr;
        const float f11 = X[11] + s11;
        X[0] = (f0 < 0.0f) ? 0.0f : f0;
        X[1] = (f1 < 0.0f) ? 0.0f : f1;
        X[2] = (f2 < 0.0f) ? 0.0f : f2;
        X[3] = (f3 < 0.0f) ? 0.0f : f3;
        X[4] = (f4 < 0.0f) ? 0.0f : f4;
        X[5] = (f5 < 0.0f) ? 0.0f : f5;
        X[6] = (f6 < 0.0f) ? 0.0f : f6;
        X[7] = (f7 < 0.0f) ? 0.0f : f7;
        X[8] = (f8 < 0.0f) ? 0.0f : f8;
        X[9] = (f9 < 0.0f) ? 0.0f : f9;
        X[10] = (f10 < 0.0f) ? 0.0f : f10;
        X[11] = (f11 < 0.0f) ? 0.0f : f11;
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_proj2obs_f32(const float *X, const double *pvd, double *obs)
{
    (void)pvd; (void)X;
    obs[0] = (double)X[8];
    obs[1] = (double)X[9];
    obs[2] = (double)X[10];
    obs[3] = (double)X[11];
}

struct plom_model_sirs4 {
    static constexpr int S = 12, N_TS = 4, INC_OFFSET = 8, N_INC = 4, PV = 48, N_CAC = 4;
    template <int V> static __device__ inline __attribute__((always_inline)) void interval(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt)
    {
        if constexpr (V == 0) plom_sirs4_interval_full_f64(X, pv, ns, t0, t1, dt, sqdt);
        else if constexpr (V == 1) plom_sirs4_interval_no_dem_sto_f64(X, pv, ns, t0, t1, dt, sqdt);
        else if constexpr (V == 2) plom_sirs4_interval_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt);
        else if constexpr (V == 3) plom_sirs4_interval_no_dem_sto_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt);
    }
    static __device__ inline __attribute__((always_inline)) void proj2obs(const double *X, const double *pvd, double *obs) { plom_sirs4_proj2obs_f64(X, pvd, obs); }
    template <int V> static __device__ inline __attribute__((always_inline)) void interval(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt)
    {
        if constexpr (V == 0) plom_sirs4_interval_full_f32(X, pv, ns, t0, t1, dt, sqdt);
        else if constexpr (V == 1) plom_sirs4_interval_no_dem_sto_f32(X, pv, ns, t0, t1, dt, sqdt);
        else if constexpr (V == 2) plom_sirs4_interval_no_env_sto_f32(X, pv, ns, t0, t1, dt, sqdt);
        else if constexpr (V == 3) plom_sirs4_interval_no_dem_sto_no_env_sto_f32(X, pv, ns, t0, t1, dt, sqdt);
    }
    static __device__ inline __attribute__((always_inline)) void proj2obs(const float *X, const double *pvd, double *obs) { plom_sirs4_proj2obs_f32(X, pvd, obs); }
    static __device__ inline __attribute__((always_inline)) double loglik(const double *obs, const double *y, const double *pvd, double t, double like_min, int *n_floor) { return plom_sirs4_loglik(obs, y, pvd, t, like_min, n_floor); }
};
#endif

#endif
