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
#define PLOM_SIRS4_N_HAT 16
#define PLOM_SIRS4_N_PP 28
#define PLOM_SIRS4_FIX_OFF 28
#define PLOM_SIRS4_N_FIX 4
#define PLOM_SIRS4_N_DRIFT 0
#define PLOM_SIRS4_DRIFT_OFFSET 8

PLOM_HD void plom_sirs4_time_funcs_f64(double tr, const double *cvk, double *out)
{
    (void)tr; (void)cvk; (void)out;
    out[0] = plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0));
    out[1] = (double)plom_recip_uniform(plom_cov_ld(cvk + 0));
    out[2] = (double)plom_recip_uniform(plom_cov_ld(cvk + 1));
    out[3] = (double)plom_recip_uniform(plom_cov_ld(cvk + 2));
    out[4] = (double)plom_recip_uniform(plom_cov_ld(cvk + 3));
}

PLOM_HD void plom_sirs4_interval_full_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const double P0 = PLOM_UNIFORM64(pv[0]);
    const double P1 = PLOM_UNIFORM64(pv[1]);
    const double P2 = PLOM_UNIFORM64(pv[2]);
    const double P3 = PLOM_UNIFORM64(pv[3]);
    const double P4 = PLOM_UNIFORM64(pv[4]);
    const double P5 = PLOM_UNIFORM64(pv[5]);
    const double P6 = PLOM_UNIFORM64(pv[6]);
    const double P7 = PLOM_UNIFORM64(pv[7]);
    const double P8 = PLOM_UNIFORM64(pv[8]);
    const double P9 = PLOM_UNIFORM64(pv[9]);
    const double P10 = PLOM_UNIFORM64(pv[10]);
    const double P11 = PLOM_UNIFORM64(pv[11]);
    const double P12 = PLOM_UNIFORM64(pv[12]);
    const double P13 = PLOM_UNIFORM64(pv[13]);
    const double P14 = PLOM_UNIFORM64(pv[14]);
    const double P15 = PLOM_UNIFORM64(pv[15]);
    const double P16 = PLOM_UNIFORM64(pv[16]);
    const double P17 = PLOM_UNIFORM64(pv[17]);
    const double P18 = PLOM_UNIFORM64(pv[18]);
    const double P19 = PLOM_UNIFORM64(pv[19]);
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P40 = PLOM_UNIFORM64(pv[40]);
    const double P41 = PLOM_UNIFORM64(pv[41]);
    const double P42 = PLOM_UNIFORM64(pv[42]);
    const double P43 = PLOM_UNIFORM64(pv[43]);
    const double P44 = PLOM_UNIFORM64(pv[44]);
    const double P45 = PLOM_UNIFORM64(pv[45]);
    const double P46 = PLOM_UNIFORM64(pv[46]);
    const double P47 = PLOM_UNIFORM64(pv[47]);
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double RC0 = tft ? tft[k * 5 + 1] : plom_recip_uniform(C0);
        const double RC1 = tft ? tft[k * 5 + 2] : plom_recip_uniform(C1);
        const double RC2 = tft ? tft[k * 5 + 3] : plom_recip_uniform(C2);
        const double RC3 = tft ? tft[k * 5 + 4] : plom_recip_uniform(C3);
        const double tf0 = tft ? tft[k * 5 + 0] : plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0));
        double s8 = 0.0;
        double s9 = 0.0;
        double s10 = 0.0;
        double s11 = 0.0;
        const double r0_0 = plom_div_rc(((P0 * (1.0 + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0, RC0);
        const double r1_0 = P8;
        const double r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        double za0, zb0;
        plom_normal_pair64(ns, &za0, &zb0);
        const double d0_0 = sqdt * za0;
        const double d1_0 = sqdt * zb0;
        double za1, zb1;
        plom_normal_pair64(ns, &za1, &zb1);
        const double d2_0 = sqdt * za1;
        const double w0_0 = sqdt * zb1;
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - plom_sqrt_rate((r0_0*X[0]))*d0_0+ plom_sqrt_rate((r2_0))*d2_0 + - (r0_0*X[0])*P16*w0_0;
        const double f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + - plom_sqrt_rate((r1_0*X[4]))*d1_0+ plom_sqrt_rate((r0_0*X[0]))*d0_0 + + (r0_0*X[0])*P16*w0_0;
        s8 += ( + (r0_0*X[0]))*dtr + + plom_sqrt_rate((r0_0*X[0]))*d0_0 + + (r0_0*X[0])*P16*w0_0;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        const double r0_1 = plom_div_rc(((P1 * (1.0 + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1, RC1);
        const double r1_1 = P9;
        const double r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        double za2, zb2;
        plom_normal_pair64(ns, &za2, &zb2);
        const double d0_1 = sqdt * za2;
        const double d1_1 = sqdt * zb2;
        double za3, zb3;
        plom_normal_pair64(ns, &za3, &zb3);
        const double d2_1 = sqdt * za3;
        const double w0_1 = sqdt * zb3;
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - plom_sqrt_rate((r0_1*X[1]))*d0_1+ plom_sqrt_rate((r2_1))*d2_1 + - (r0_1*X[1])*P17*w0_1;
        const double f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + - plom_sqrt_rate((r1_1*X[5]))*d1_1+ plom_sqrt_rate((r0_1*X[1]))*d0_1 + + (r0_1*X[1])*P17*w0_1;
        s9 += ( + (r0_1*X[1]))*dtr + + plom_sqrt_rate((r0_1*X[1]))*d0_1 + + (r0_1*X[1])*P17*w0_1;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        const double r0_2 = plom_div_rc(((P2 * (1.0 + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2, RC2);
        const double r1_2 = P10;
        const double r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        double za4, zb4;
        plom_normal_pair64(ns, &za4, &zb4);
        const double d0_2 = sqdt * za4;
        const double d1_2 = sqdt * zb4;
        double za5, zb5;
        plom_normal_pair64(ns, &za5, &zb5);
        const double d2_2 = sqdt * za5;
        const double w0_2 = sqdt * zb5;
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - plom_sqrt_rate((r0_2*X[2]))*d0_2+ plom_sqrt_rate((r2_2))*d2_2 + - (r0_2*X[2])*P18*w0_2;
        const double f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + - plom_sqrt_rate((r1_2*X[6]))*d1_2+ plom_sqrt_rate((r0_2*X[2]))*d0_2 + + (r0_2*X[2])*P18*w0_2;
        s10 += ( + (r0_2*X[2]))*dtr + + plom_sqrt_rate((r0_2*X[2]))*d0_2 + + (r0_2*X[2])*P18*w0_2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        const double r0_3 = plom_div_rc(((P3 * (1.0 + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3, RC3);
        const double r1_3 = P11;
        const double r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        double za6, zb6;
        plom_normal_pair64(ns, &za6, &zb6);
        const double d0_3 = sqdt * za6;
        const double d1_3 = sqdt * zb6;
        double za7, zb7;
        plom_normal_pair64(ns, &za7, &zb7);
        const double d2_3 = sqdt * za7;
        const double w0_3 = sqdt * zb7;
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - plom_sqrt_rate((r0_3*X[3]))*d0_3+ plom_sqrt_rate((r2_3))*d2_3 + - (r0_3*X[3])*P19*w0_3;
        const double f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + - plom_sqrt_rate((r1_3*X[7]))*d1_3+ plom_sqrt_rate((r0_3*X[3]))*d0_3 + + (r0_3*X[3])*P19*w0_3;
        s11 += ( + (r0_3*X[3]))*dtr + + plom_sqrt_rate((r0_3*X[3]))*d0_3 + + (r0_3*X[3])*P19*w0_3;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        const double f8 = X[8] + s8;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        const double f9 = X[9] + s9;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        const double f10 = X[10] + s10;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        const double f11 = X[11] + s11;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        t += dt;
    }
}

PLOM_HD void plom_sirs4_interval_no_dem_sto_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const double P0 = PLOM_UNIFORM64(pv[0]);
    const double P1 = PLOM_UNIFORM64(pv[1]);
    const double P2 = PLOM_UNIFORM64(pv[2]);
    const double P3 = PLOM_UNIFORM64(pv[3]);
    const double P4 = PLOM_UNIFORM64(pv[4]);
    const double P5 = PLOM_UNIFORM64(pv[5]);
    const double P6 = PLOM_UNIFORM64(pv[6]);
    const double P7 = PLOM_UNIFORM64(pv[7]);
    const double P8 = PLOM_UNIFORM64(pv[8]);
    const double P9 = PLOM_UNIFORM64(pv[9]);
    const double P10 = PLOM_UNIFORM64(pv[10]);
    const double P11 = PLOM_UNIFORM64(pv[11]);
    const double P12 = PLOM_UNIFORM64(pv[12]);
    const double P13 = PLOM_UNIFORM64(pv[13]);
    const double P14 = PLOM_UNIFORM64(pv[14]);
    const double P15 = PLOM_UNIFORM64(pv[15]);
    const double P16 = PLOM_UNIFORM64(pv[16]);
    const double P17 = PLOM_UNIFORM64(pv[17]);
    const double P18 = PLOM_UNIFORM64(pv[18]);
    const double P19 = PLOM_UNIFORM64(pv[19]);
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P40 = PLOM_UNIFORM64(pv[40]);
    const double P41 = PLOM_UNIFORM64(pv[41]);
    const double P42 = PLOM_UNIFORM64(pv[42]);
    const double P43 = PLOM_UNIFORM64(pv[43]);
    const double P44 = PLOM_UNIFORM64(pv[44]);
    const double P45 = PLOM_UNIFORM64(pv[45]);
    const double P46 = PLOM_UNIFORM64(pv[46]);
    const double P47 = PLOM_UNIFORM64(pv[47]);
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double RC0 = tft ? tft[k * 5 + 1] : plom_recip_uniform(C0);
        const double RC1 = tft ? tft[k * 5 + 2] : plom_recip_uniform(C1);
        const double RC2 = tft ? tft[k * 5 + 3] : plom_recip_uniform(C2);
        const double RC3 = tft ? tft[k * 5 + 4] : plom_recip_uniform(C3);
        const double tf0 = tft ? tft[k * 5 + 0] : plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0));
        double s8 = 0.0;
        double s9 = 0.0;
        double s10 = 0.0;
        double s11 = 0.0;
        const double r0_0 = plom_div_rc(((P0 * (1.0 + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0, RC0);
        const double r1_0 = P8;
        const double r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        double za0, zb0;
        plom_normal_pair64(ns, &za0, &zb0);
        const double w0_0 = sqdt * za0;
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - (r0_0*X[0])*P16*w0_0;
        const double f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + + (r0_0*X[0])*P16*w0_0;
        s8 += ( + (r0_0*X[0]))*dtr + + (r0_0*X[0])*P16*w0_0;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        const double r0_1 = plom_div_rc(((P1 * (1.0 + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1, RC1);
        const double r1_1 = P9;
        const double r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        const double w0_1 = sqdt * zb0;
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - (r0_1*X[1])*P17*w0_1;
        const double f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + + (r0_1*X[1])*P17*w0_1;
        s9 += ( + (r0_1*X[1]))*dtr + + (r0_1*X[1])*P17*w0_1;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        const double r0_2 = plom_div_rc(((P2 * (1.0 + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2, RC2);
        const double r1_2 = P10;
        const double r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        double za1, zb1;
        plom_normal_pair64(ns, &za1, &zb1);
        const double w0_2 = sqdt * za1;
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - (r0_2*X[2])*P18*w0_2;
        const double f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + + (r0_2*X[2])*P18*w0_2;
        s10 += ( + (r0_2*X[2]))*dtr + + (r0_2*X[2])*P18*w0_2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        const double r0_3 = plom_div_rc(((P3 * (1.0 + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3, RC3);
        const double r1_3 = P11;
        const double r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        const double w0_3 = sqdt * zb1;
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - (r0_3*X[3])*P19*w0_3;
        const double f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + + (r0_3*X[3])*P19*w0_3;
        s11 += ( + (r0_3*X[3]))*dtr + + (r0_3*X[3])*P19*w0_3;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        const double f8 = X[8] + s8;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        const double f9 = X[9] + s9;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        const double f10 = X[10] + s10;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        const double f11 = X[11] + s11;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        t += dt;
    }
}

PLOM_HD void plom_sirs4_interval_no_env_sto_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const double P0 = PLOM_UNIFORM64(pv[0]);
    const double P1 = PLOM_UNIFORM64(pv[1]);
    const double P2 = PLOM_UNIFORM64(pv[2]);
    const double P3 = PLOM_UNIFORM64(pv[3]);
    const double P4 = PLOM_UNIFORM64(pv[4]);
    const double P5 = PLOM_UNIFORM64(pv[5]);
    const double P6 = PLOM_UNIFORM64(pv[6]);
    const double P7 = PLOM_UNIFORM64(pv[7]);
    const double P8 = PLOM_UNIFORM64(pv[8]);
    const double P9 = PLOM_UNIFORM64(pv[9]);
    const double P10 = PLOM_UNIFORM64(pv[10]);
    const double P11 = PLOM_UNIFORM64(pv[11]);
    const double P12 = PLOM_UNIFORM64(pv[12]);
    const double P13 = PLOM_UNIFORM64(pv[13]);
    const double P14 = PLOM_UNIFORM64(pv[14]);
    const double P15 = PLOM_UNIFORM64(pv[15]);
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P40 = PLOM_UNIFORM64(pv[40]);
    const double P41 = PLOM_UNIFORM64(pv[41]);
    const double P42 = PLOM_UNIFORM64(pv[42]);
    const double P43 = PLOM_UNIFORM64(pv[43]);
    const double P44 = PLOM_UNIFORM64(pv[44]);
    const double P45 = PLOM_UNIFORM64(pv[45]);
    const double P46 = PLOM_UNIFORM64(pv[46]);
    const double P47 = PLOM_UNIFORM64(pv[47]);
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double RC0 = tft ? tft[k * 5 + 1] : plom_recip_uniform(C0);
        const double RC1 = tft ? tft[k * 5 + 2] : plom_recip_uniform(C1);
        const double RC2 = tft ? tft[k * 5 + 3] : plom_recip_uniform(C2);
        const double RC3 = tft ? tft[k * 5 + 4] : plom_recip_uniform(C3);
        const double tf0 = tft ? tft[k * 5 + 0] : plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0));
        double s8 = 0.0;
        double s9 = 0.0;
        double s10 = 0.0;
        double s11 = 0.0;
        const double r0_0 = plom_div_rc(((P0 * (1.0 + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0, RC0);
        const double r1_0 = P8;
        const double r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        double za0, zb0;
        plom_normal_pair64(ns, &za0, &zb0);
        const double d0_0 = sqdt * za0;
        const double d1_0 = sqdt * zb0;
        double za1, zb1;
        plom_normal_pair64(ns, &za1, &zb1);
        const double d2_0 = sqdt * za1;
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - plom_sqrt_rate((r0_0*X[0]))*d0_0+ plom_sqrt_rate((r2_0))*d2_0;
        const double f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + - plom_sqrt_rate((r1_0*X[4]))*d1_0+ plom_sqrt_rate((r0_0*X[0]))*d0_0;
        s8 += ( + (r0_0*X[0]))*dtr + + plom_sqrt_rate((r0_0*X[0]))*d0_0;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        const double r0_1 = plom_div_rc(((P1 * (1.0 + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1, RC1);
        const double r1_1 = P9;
        const double r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        const double d0_1 = sqdt * zb1;
        double za2, zb2;
        plom_normal_pair64(ns, &za2, &zb2);
        const double d1_1 = sqdt * za2;
        const double d2_1 = sqdt * zb2;
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - plom_sqrt_rate((r0_1*X[1]))*d0_1+ plom_sqrt_rate((r2_1))*d2_1;
        const double f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + - plom_sqrt_rate((r1_1*X[5]))*d1_1+ plom_sqrt_rate((r0_1*X[1]))*d0_1;
        s9 += ( + (r0_1*X[1]))*dtr + + plom_sqrt_rate((r0_1*X[1]))*d0_1;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        const double r0_2 = plom_div_rc(((P2 * (1.0 + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2, RC2);
        const double r1_2 = P10;
        const double r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        double za3, zb3;
        plom_normal_pair64(ns, &za3, &zb3);
        const double d0_2 = sqdt * za3;
        const double d1_2 = sqdt * zb3;
        double za4, zb4;
        plom_normal_pair64(ns, &za4, &zb4);
        const double d2_2 = sqdt * za4;
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - plom_sqrt_rate((r0_2*X[2]))*d0_2+ plom_sqrt_rate((r2_2))*d2_2;
        const double f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + - plom_sqrt_rate((r1_2*X[6]))*d1_2+ plom_sqrt_rate((r0_2*X[2]))*d0_2;
        s10 += ( + (r0_2*X[2]))*dtr + + plom_sqrt_rate((r0_2*X[2]))*d0_2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        const double r0_3 = plom_div_rc(((P3 * (1.0 + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3, RC3);
        const double r1_3 = P11;
        const double r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        const double d0_3 = sqdt * zb4;
        double za5, zb5;
        plom_normal_pair64(ns, &za5, &zb5);
        const double d1_3 = sqdt * za5;
        const double d2_3 = sqdt * zb5;
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - plom_sqrt_rate((r0_3*X[3]))*d0_3+ plom_sqrt_rate((r2_3))*d2_3;
        const double f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + - plom_sqrt_rate((r1_3*X[7]))*d1_3+ plom_sqrt_rate((r0_3*X[3]))*d0_3;
        s11 += ( + (r0_3*X[3]))*dtr + + plom_sqrt_rate((r0_3*X[3]))*d0_3;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        const double f8 = X[8] + s8;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        const double f9 = X[9] + s9;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        const double f10 = X[10] + s10;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        const double f11 = X[11] + s11;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        t += dt;
    }
}

PLOM_HD void plom_sirs4_interval_no_dem_sto_no_env_sto_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const double P0 = PLOM_UNIFORM64(pv[0]);
    const double P1 = PLOM_UNIFORM64(pv[1]);
    const double P2 = PLOM_UNIFORM64(pv[2]);
    const double P3 = PLOM_UNIFORM64(pv[3]);
    const double P4 = PLOM_UNIFORM64(pv[4]);
    const double P5 = PLOM_UNIFORM64(pv[5]);
    const double P6 = PLOM_UNIFORM64(pv[6]);
    const double P7 = PLOM_UNIFORM64(pv[7]);
    const double P8 = PLOM_UNIFORM64(pv[8]);
    const double P9 = PLOM_UNIFORM64(pv[9]);
    const double P10 = PLOM_UNIFORM64(pv[10]);
    const double P11 = PLOM_UNIFORM64(pv[11]);
    const double P12 = PLOM_UNIFORM64(pv[12]);
    const double P13 = PLOM_UNIFORM64(pv[13]);
    const double P14 = PLOM_UNIFORM64(pv[14]);
    const double P15 = PLOM_UNIFORM64(pv[15]);
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P40 = PLOM_UNIFORM64(pv[40]);
    const double P41 = PLOM_UNIFORM64(pv[41]);
    const double P42 = PLOM_UNIFORM64(pv[42]);
    const double P43 = PLOM_UNIFORM64(pv[43]);
    const double P44 = PLOM_UNIFORM64(pv[44]);
    const double P45 = PLOM_UNIFORM64(pv[45]);
    const double P46 = PLOM_UNIFORM64(pv[46]);
    const double P47 = PLOM_UNIFORM64(pv[47]);
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double RC0 = tft ? tft[k * 5 + 1] : plom_recip_uniform(C0);
        const double RC1 = tft ? tft[k * 5 + 2] : plom_recip_uniform(C1);
        const double RC2 = tft ? tft[k * 5 + 3] : plom_recip_uniform(C2);
        const double RC3 = tft ? tft[k * 5 + 4] : plom_recip_uniform(C3);
        const double tf0 = tft ? tft[k * 5 + 0] : plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0));
        double s8 = 0.0;
        double s9 = 0.0;
        double s10 = 0.0;
        double s11 = 0.0;
        const double r0_0 = plom_div_rc(((P0 * (1.0 + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0, RC0);
        const double r1_0 = P8;
        const double r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr;
        const double f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr;
        s8 += ( + (r0_0*X[0]))*dtr;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        const double r0_1 = plom_div_rc(((P1 * (1.0 + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1, RC1);
        const double r1_1 = P9;
        const double r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr;
        const double f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr;
        s9 += ( + (r0_1*X[1]))*dtr;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        const double r0_2 = plom_div_rc(((P2 * (1.0 + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2, RC2);
        const double r1_2 = P10;
        const double r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr;
        const double f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr;
        s10 += ( + (r0_2*X[2]))*dtr;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        const double r0_3 = plom_div_rc(((P3 * (1.0 + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3, RC3);
        const double r1_3 = P11;
        const double r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr;
        const double f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr;
        s11 += ( + (r0_3*X[3]))*dtr;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        const double f8 = X[8] + s8;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        const double f9 = X[9] + s9;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        const double f10 = X[10] + s10;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        const double f11 = X[11] + s11;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        t += dt;
    }
}

PLOM_HD void plom_sirs4_proj2obs_f64(const double *X, const double *pvd, const double *cvd, double *obs)
{
    (void)pvd; (void)cvd; (void)X;
    obs[0] = (double)X[8];
    obs[1] = (double)X[9];
    obs[2] = (double)X[10];
    obs[3] = (double)X[11];
}

PLOM_HD double plom_sirs4_loglik(const double *obs, const double *y, const double *pvd, const double *cvd, double t, double like_min, int *n_floor)
{
    double loglike = 0.0;
    int nf = 0;
    (void)t; (void)pvd; (void)cvd;
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

PLOM_HD void plom_sirs4_obs_moments(const double *obs, const double *pvd, const double *cvd, double t, double *mean, double *var)
{
    (void)t; (void)pvd; (void)cvd;
    {
        const double x = obs[0];
        mean[0] = (pvd[20] * x);
        var[0] = (((pvd[20] * (1.0 - pvd[20])) * x) + plom_powi(((pvd[20] * pvd[24]) * x), 2));
    }
    {
        const double x = obs[1];
        mean[1] = (pvd[21] * x);
        var[1] = (((pvd[21] * (1.0 - pvd[21])) * x) + plom_powi(((pvd[21] * pvd[25]) * x), 2));
    }
    {
        const double x = obs[2];
        mean[2] = (pvd[22] * x);
        var[2] = (((pvd[22] * (1.0 - pvd[22])) * x) + plom_powi(((pvd[22] * pvd[26]) * x), 2));
    }
    {
        const double x = obs[3];
        mean[3] = (pvd[23] * x);
        var[3] = (((pvd[23] * (1.0 - pvd[23])) * x) + plom_powi(((pvd[23] * pvd[27]) * x), 2));
    }
}

PLOM_HD void plom_sirs4_hat_values(const double *X, const double *pvd, const double *cvd, const plom_drift_view *dm, double t, double *out)
{
    (void)t; (void)pvd; (void)cvd; (void)dm;
    out[0] = X[0];
    out[1] = X[1];
    out[2] = X[2];
    out[3] = X[3];
    out[4] = X[4];
    out[5] = X[5];
    out[6] = X[6];
    out[7] = X[7];
    out[8] = cvd[0] - (X[0] + X[4]);
    out[9] = cvd[1] - (X[1] + X[5]);
    out[10] = cvd[2] - (X[2] + X[6]);
    out[11] = cvd[3] - (X[3] + X[7]);
    double obs[4], var[4];
    plom_sirs4_proj2obs_f64(X, pvd, cvd, obs);
    plom_sirs4_obs_moments(obs, pvd, cvd, t, out + 12, var);
}

PLOM_HD void plom_sirs4_interval_psr_full_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const double P0 = PLOM_UNIFORM64(pv[0]);
    const double P1 = PLOM_UNIFORM64(pv[1]);
    const double P2 = PLOM_UNIFORM64(pv[2]);
    const double P3 = PLOM_UNIFORM64(pv[3]);
    const double P4 = PLOM_UNIFORM64(pv[4]);
    const double P5 = PLOM_UNIFORM64(pv[5]);
    const double P6 = PLOM_UNIFORM64(pv[6]);
    const double P7 = PLOM_UNIFORM64(pv[7]);
    const double P8 = PLOM_UNIFORM64(pv[8]);
    const double P9 = PLOM_UNIFORM64(pv[9]);
    const double P10 = PLOM_UNIFORM64(pv[10]);
    const double P11 = PLOM_UNIFORM64(pv[11]);
    const double P12 = PLOM_UNIFORM64(pv[12]);
    const double P13 = PLOM_UNIFORM64(pv[13]);
    const double P14 = PLOM_UNIFORM64(pv[14]);
    const double P15 = PLOM_UNIFORM64(pv[15]);
    const double P16 = PLOM_UNIFORM64(pv[16]);
    const double P17 = PLOM_UNIFORM64(pv[17]);
    const double P18 = PLOM_UNIFORM64(pv[18]);
    const double P19 = PLOM_UNIFORM64(pv[19]);
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P40 = PLOM_UNIFORM64(pv[40]);
    const double P41 = PLOM_UNIFORM64(pv[41]);
    const double P42 = PLOM_UNIFORM64(pv[42]);
    const double P43 = PLOM_UNIFORM64(pv[43]);
    const double P44 = PLOM_UNIFORM64(pv[44]);
    const double P45 = PLOM_UNIFORM64(pv[45]);
    const double P46 = PLOM_UNIFORM64(pv[46]);
    const double P47 = PLOM_UNIFORM64(pv[47]);
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double RC0 = tft ? tft[k * 5 + 1] : plom_recip_uniform(C0);
        const double RC1 = tft ? tft[k * 5 + 2] : plom_recip_uniform(C1);
        const double RC2 = tft ? tft[k * 5 + 3] : plom_recip_uniform(C2);
        const double RC3 = tft ? tft[k * 5 + 4] : plom_recip_uniform(C3);
        const double tf0 = tft ? tft[k * 5 + 0] : plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0));
        double n0_0 = 0.0;
        double n1_0 = 0.0;
        double n2_0 = 0.0;
        double n0_1 = 0.0;
        double n1_1 = 0.0;
        double n2_1 = 0.0;
        double n0_2 = 0.0;
        double n1_2 = 0.0;
        double n2_2 = 0.0;
        double n0_3 = 0.0;
        double n1_3 = 0.0;
        double n2_3 = 0.0;
        double stay0 = 0.0;
        double stay1 = 0.0;
        double stay2 = 0.0;
        double stay3 = 0.0;
        double stay4 = 0.0;
        double stay5 = 0.0;
        double stay6 = 0.0;
        double stay7 = 0.0;
        (void)n0_0; (void)n1_0; (void)n2_0; (void)n0_1; (void)n1_1; (void)n2_1; (void)n0_2; (void)n1_2; (void)n2_2; (void)n0_3; (void)n1_3; (void)n2_3;
        const double r0_0 = plom_div_rc(((P0 * (1.0 + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0, RC0);
        const double r1_0 = P8;
        const double r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        const double r0_1 = plom_div_rc(((P1 * (1.0 + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1, RC1);
        const double r1_1 = P9;
        const double r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        const double r0_2 = plom_div_rc(((P2 * (1.0 + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2, RC2);
        const double r1_2 = P10;
        const double r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        const double r0_3 = plom_div_rc(((P3 * (1.0 + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3, RC3);
        const double r1_3 = P11;
        const double r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        const double sd0_0 = P16;
        const double v0_0 = sd0_0 * sd0_0;
        const double xi0_0 = (r0_0 > 0.0) ? (v0_0 * plom_gamma_draw(ns, dt / v0_0)) / dt : 1.0;
        {
            const double sum = (r0_0 * xi0_0) * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * (((r0_0 * xi0_0) * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[0]);
            n0_0 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay0 = (p0 == p0 && cnt > 0.0) ? cnt - n0_0 : 0.0;
        }
        {
            const double sum = r1_0 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_0 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[4]);
            n1_0 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay4 = (p0 == p0 && cnt > 0.0) ? cnt - n1_0 : 0.0;
        }
        {
            const double mu = r2_0 * dt;
            n2_0 = (mu > 0.0) ? plom_poisson_draw(ns, mu) : 0.0;
        }
        const double sd0_1 = P17;
        const double v0_1 = sd0_1 * sd0_1;
        const double xi0_1 = (r0_1 > 0.0) ? (v0_1 * plom_gamma_draw(ns, dt / v0_1)) / dt : 1.0;
        {
            const double sum = (r0_1 * xi0_1) * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * (((r0_1 * xi0_1) * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[1]);
            n0_1 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay1 = (p0 == p0 && cnt > 0.0) ? cnt - n0_1 : 0.0;
        }
        {
            const double sum = r1_1 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_1 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[5]);
            n1_1 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay5 = (p0 == p0 && cnt > 0.0) ? cnt - n1_1 : 0.0;
        }
        {
            const double mu = r2_1 * dt;
            n2_1 = (mu > 0.0) ? plom_poisson_draw(ns, mu) : 0.0;
        }
        const double sd0_2 = P18;
        const double v0_2 = sd0_2 * sd0_2;
        const double xi0_2 = (r0_2 > 0.0) ? (v0_2 * plom_gamma_draw(ns, dt / v0_2)) / dt : 1.0;
        {
            const double sum = (r0_2 * xi0_2) * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * (((r0_2 * xi0_2) * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[2]);
            n0_2 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay2 = (p0 == p0 && cnt > 0.0) ? cnt - n0_2 : 0.0;
        }
        {
            const double sum = r1_2 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_2 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[6]);
            n1_2 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay6 = (p0 == p0 && cnt > 0.0) ? cnt - n1_2 : 0.0;
        }
        {
            const double mu = r2_2 * dt;
            n2_2 = (mu > 0.0) ? plom_poisson_draw(ns, mu) : 0.0;
        }
        const double sd0_3 = P19;
        const double v0_3 = sd0_3 * sd0_3;
        const double xi0_3 = (r0_3 > 0.0) ? (v0_3 * plom_gamma_draw(ns, dt / v0_3)) / dt : 1.0;
        {
            const double sum = (r0_3 * xi0_3) * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * (((r0_3 * xi0_3) * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[3]);
            n0_3 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay3 = (p0 == p0 && cnt > 0.0) ? cnt - n0_3 : 0.0;
        }
        {
            const double sum = r1_3 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_3 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[7]);
            n1_3 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay7 = (p0 == p0 && cnt > 0.0) ? cnt - n1_3 : 0.0;
        }
        {
            const double mu = r2_3 * dt;
            n2_3 = (mu > 0.0) ? plom_poisson_draw(ns, mu) : 0.0;
        }
        X[0] = stay0 + n2_0;
        X[4] = stay4 + n0_0;
        X[1] = stay1 + n2_1;
        X[5] = stay5 + n0_1;
        X[2] = stay2 + n2_2;
        X[6] = stay6 + n0_2;
        X[3] = stay3 + n2_3;
        X[7] = stay7 + n0_3;
        X[8] = X[8] + (n0_0);
        X[9] = X[9] + (n0_1);
        X[10] = X[10] + (n0_2);
        X[11] = X[11] + (n0_3);
        t += dt;
    }
}

PLOM_HD void plom_sirs4_interval_psr_no_env_sto_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const double P0 = PLOM_UNIFORM64(pv[0]);
    const double P1 = PLOM_UNIFORM64(pv[1]);
    const double P2 = PLOM_UNIFORM64(pv[2]);
    const double P3 = PLOM_UNIFORM64(pv[3]);
    const double P4 = PLOM_UNIFORM64(pv[4]);
    const double P5 = PLOM_UNIFORM64(pv[5]);
    const double P6 = PLOM_UNIFORM64(pv[6]);
    const double P7 = PLOM_UNIFORM64(pv[7]);
    const double P8 = PLOM_UNIFORM64(pv[8]);
    const double P9 = PLOM_UNIFORM64(pv[9]);
    const double P10 = PLOM_UNIFORM64(pv[10]);
    const double P11 = PLOM_UNIFORM64(pv[11]);
    const double P12 = PLOM_UNIFORM64(pv[12]);
    const double P13 = PLOM_UNIFORM64(pv[13]);
    const double P14 = PLOM_UNIFORM64(pv[14]);
    const double P15 = PLOM_UNIFORM64(pv[15]);
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P40 = PLOM_UNIFORM64(pv[40]);
    const double P41 = PLOM_UNIFORM64(pv[41]);
    const double P42 = PLOM_UNIFORM64(pv[42]);
    const double P43 = PLOM_UNIFORM64(pv[43]);
    const double P44 = PLOM_UNIFORM64(pv[44]);
    const double P45 = PLOM_UNIFORM64(pv[45]);
    const double P46 = PLOM_UNIFORM64(pv[46]);
    const double P47 = PLOM_UNIFORM64(pv[47]);
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double RC0 = tft ? tft[k * 5 + 1] : plom_recip_uniform(C0);
        const double RC1 = tft ? tft[k * 5 + 2] : plom_recip_uniform(C1);
        const double RC2 = tft ? tft[k * 5 + 3] : plom_recip_uniform(C2);
        const double RC3 = tft ? tft[k * 5 + 4] : plom_recip_uniform(C3);
        const double tf0 = tft ? tft[k * 5 + 0] : plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0));
        double n0_0 = 0.0;
        double n1_0 = 0.0;
        double n2_0 = 0.0;
        double n0_1 = 0.0;
        double n1_1 = 0.0;
        double n2_1 = 0.0;
        double n0_2 = 0.0;
        double n1_2 = 0.0;
        double n2_2 = 0.0;
        double n0_3 = 0.0;
        double n1_3 = 0.0;
        double n2_3 = 0.0;
        double stay0 = 0.0;
        double stay1 = 0.0;
        double stay2 = 0.0;
        double stay3 = 0.0;
        double stay4 = 0.0;
        double stay5 = 0.0;
        double stay6 = 0.0;
        double stay7 = 0.0;
        (void)n0_0; (void)n1_0; (void)n2_0; (void)n0_1; (void)n1_1; (void)n2_1; (void)n0_2; (void)n1_2; (void)n2_2; (void)n0_3; (void)n1_3; (void)n2_3;
        const double r0_0 = plom_div_rc(((P0 * (1.0 + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0, RC0);
        const double r1_0 = P8;
        const double r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        const double r0_1 = plom_div_rc(((P1 * (1.0 + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1, RC1);
        const double r1_1 = P9;
        const double r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        const double r0_2 = plom_div_rc(((P2 * (1.0 + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2, RC2);
        const double r1_2 = P10;
        const double r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        const double r0_3 = plom_div_rc(((P3 * (1.0 + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3, RC3);
        const double r1_3 = P11;
        const double r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        const double xi0_0 = 1.0;
        {
            const double sum = (r0_0 * xi0_0) * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * (((r0_0 * xi0_0) * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[0]);
            n0_0 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay0 = (p0 == p0 && cnt > 0.0) ? cnt - n0_0 : 0.0;
        }
        {
            const double sum = r1_0 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_0 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[4]);
            n1_0 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay4 = (p0 == p0 && cnt > 0.0) ? cnt - n1_0 : 0.0;
        }
        {
            const double mu = r2_0 * dt;
            n2_0 = (mu > 0.0) ? plom_poisson_draw(ns, mu) : 0.0;
        }
        const double xi0_1 = 1.0;
        {
            const double sum = (r0_1 * xi0_1) * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * (((r0_1 * xi0_1) * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[1]);
            n0_1 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay1 = (p0 == p0 && cnt > 0.0) ? cnt - n0_1 : 0.0;
        }
        {
            const double sum = r1_1 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_1 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[5]);
            n1_1 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay5 = (p0 == p0 && cnt > 0.0) ? cnt - n1_1 : 0.0;
        }
        {
            const double mu = r2_1 * dt;
            n2_1 = (mu > 0.0) ? plom_poisson_draw(ns, mu) : 0.0;
        }
        const double xi0_2 = 1.0;
        {
            const double sum = (r0_2 * xi0_2) * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * (((r0_2 * xi0_2) * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[2]);
            n0_2 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay2 = (p0 == p0 && cnt > 0.0) ? cnt - n0_2 : 0.0;
        }
        {
            const double sum = r1_2 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_2 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[6]);
            n1_2 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay6 = (p0 == p0 && cnt > 0.0) ? cnt - n1_2 : 0.0;
        }
        {
            const double mu = r2_2 * dt;
            n2_2 = (mu > 0.0) ? plom_poisson_draw(ns, mu) : 0.0;
        }
        const double xi0_3 = 1.0;
        {
            const double sum = (r0_3 * xi0_3) * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * (((r0_3 * xi0_3) * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[3]);
            n0_3 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay3 = (p0 == p0 && cnt > 0.0) ? cnt - n0_3 : 0.0;
        }
        {
            const double sum = r1_3 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_3 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[7]);
            n1_3 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay7 = (p0 == p0 && cnt > 0.0) ? cnt - n1_3 : 0.0;
        }
        {
            const double mu = r2_3 * dt;
            n2_3 = (mu > 0.0) ? plom_poisson_draw(ns, mu) : 0.0;
        }
        X[0] = stay0 + n2_0;
        X[4] = stay4 + n0_0;
        X[1] = stay1 + n2_1;
        X[5] = stay5 + n0_1;
        X[2] = stay2 + n2_2;
        X[6] = stay6 + n0_2;
        X[3] = stay3 + n2_3;
        X[7] = stay7 + n0_3;
        X[8] = X[8] + (n0_0);
        X[9] = X[9] + (n0_1);
        X[10] = X[10] + (n0_2);
        X[11] = X[11] + (n0_3);
        t += dt;
    }
}

PLOM_HD void plom_sirs4_ode_rhs(double t, const double *X, const double *pv, const double *lp, const double *cv, const plom_spline_set *sp, double *f)
{
    (void)pv; (void)lp; (void)cv; (void)sp;
    const double tr = t; (void)tr;
    const double P0 = PLOM_UNIFORM64(pv[0]);
    const double P1 = PLOM_UNIFORM64(pv[1]);
    const double P2 = PLOM_UNIFORM64(pv[2]);
    const double P3 = PLOM_UNIFORM64(pv[3]);
    const double P4 = PLOM_UNIFORM64(pv[4]);
    const double P5 = PLOM_UNIFORM64(pv[5]);
    const double P6 = PLOM_UNIFORM64(pv[6]);
    const double P7 = PLOM_UNIFORM64(pv[7]);
    const double P8 = PLOM_UNIFORM64(pv[8]);
    const double P9 = PLOM_UNIFORM64(pv[9]);
    const double P10 = PLOM_UNIFORM64(pv[10]);
    const double P11 = PLOM_UNIFORM64(pv[11]);
    const double P12 = PLOM_UNIFORM64(pv[12]);
    const double P13 = PLOM_UNIFORM64(pv[13]);
    const double P14 = PLOM_UNIFORM64(pv[14]);
    const double P15 = PLOM_UNIFORM64(pv[15]);
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P40 = PLOM_UNIFORM64(pv[40]);
    const double P41 = PLOM_UNIFORM64(pv[41]);
    const double P42 = PLOM_UNIFORM64(pv[42]);
    const double P43 = PLOM_UNIFORM64(pv[43]);
    const double P44 = PLOM_UNIFORM64(pv[44]);
    const double P45 = PLOM_UNIFORM64(pv[45]);
    const double P46 = PLOM_UNIFORM64(pv[46]);
    const double P47 = PLOM_UNIFORM64(pv[47]);
    const double C0 = sp ? plom_spline_at(sp, 0, t) : plom_cov_ld(cv + 0);
    const double C1 = sp ? plom_spline_at(sp, 1, t) : plom_cov_ld(cv + 1);
    const double C2 = sp ? plom_spline_at(sp, 2, t) : plom_cov_ld(cv + 2);
    const double C3 = sp ? plom_spline_at(sp, 3, t) : plom_cov_ld(cv + 3);
    const double tf0 = plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0));
    const double r0_0 = (((P0 * (1.0 + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])) / C0);
    const double r1_0 = P8;
    const double r2_0 = (P12 * ((C0 - X[0]) - X[4]));
    const double r0_1 = (((P1 * (1.0 + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])) / C1);
    const double r1_1 = P9;
    const double r2_1 = (P13 * ((C1 - X[1]) - X[5]));
    const double r0_2 = (((P2 * (1.0 + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])) / C2);
    const double r1_2 = P10;
    const double r2_2 = (P14 * ((C2 - X[2]) - X[6]));
    const double r0_3 = (((P3 * (1.0 + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])) / C3);
    const double r1_3 = P11;
    const double r2_3 = (P15 * ((C3 - X[3]) - X[7]));
    f[0] =  - (r0_0*X[0]) + (r2_0);
    f[4] =  - (r1_0*X[4]) + (r0_0*X[0]);
    f[1] =  - (r0_1*X[1]) + (r2_1);
    f[5] =  - (r1_1*X[5]) + (r0_1*X[1]);
    f[2] =  - (r0_2*X[2]) + (r2_2);
    f[6] =  - (r1_2*X[6]) + (r0_2*X[2]);
    f[3] =  - (r0_3*X[3]) + (r2_3);
    f[7] =  - (r1_3*X[7]) + (r0_3*X[3]);
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r0_0*X[0]);
        f[8] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r0_1*X[1]);
        f[9] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r0_2*X[2]);
        f[10] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r0_3*X[3]);
        f[11] = sum_inc;
    }
}

#define PLOM_ODE_FN plom_sirs4_ode_integrate
#define PLOM_ODE_RHS plom_sirs4_ode_rhs
#define PLOM_ODE_N 12
#define PLOM_ODE_QUAL PLOM_HD
#include "plom_ode_impl.h"

PLOM_HD int plom_sirs4_interval_ode_f64(double *X, const double *pv, const double *cv, const plom_spline_set *sp, double t0, double t1, double dt, double eps_abs, double eps_rel)
{
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    return plom_sirs4_ode_integrate(X, pv, 0, cv, sp, t0, t1, dt, eps_abs, eps_rel);
}

PLOM_HD void plom_sirs4_interval_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, unsigned int noises_off, const double *cv, int cs, double eps_abs, double eps_rel, const plom_drift_view *dm, const plom_spline_set *sp)
{
    if (noises_off & PLOM_IMPL_ODE) {
        (void)plom_sirs4_interval_ode_f64(X, pv, cv, sp, t0, t1, dt, eps_abs, eps_rel);
        return;
    }
    if (noises_off & PLOM_IMPL_PSR) {
        if (noises_off & PLOM_NO_ENV_STO) plom_sirs4_interval_psr_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
        else plom_sirs4_interval_psr_full_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
        return;
    }
    const unsigned int v = noises_off & (PLOM_NO_DEM_STO | PLOM_NO_ENV_STO);
    if (v == 0u) plom_sirs4_interval_full_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
    else if (v == PLOM_NO_DEM_STO) plom_sirs4_interval_no_dem_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
    else if (v == PLOM_NO_ENV_STO) plom_sirs4_interval_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
    else plom_sirs4_interval_no_dem_sto_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
}

#if defined(__HIPCC__)
static __device__ inline __attribute__((always_inline)) void plom_sirs4_time_funcs_f32(float tr, const double *cvk, float *out)
{
    (void)tr; (void)cvk; (void)out;
    out[0] = plom_sinf(plom_fdivf(((2.0f * 3.141592653589793f) * tr), 365.0f));
    out[1] = (float)plom_recip_uniform(plom_cov_ld(cvk + 0));
    out[2] = (float)plom_recip_uniform(plom_cov_ld(cvk + 1));
    out[3] = (float)plom_recip_uniform(plom_cov_ld(cvk + 2));
    out[4] = (float)plom_recip_uniform(plom_cov_ld(cvk + 3));
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_full_f32(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const float P0 = PLOM_UNIFORM32(pv[0]);
    const float P1 = PLOM_UNIFORM32(pv[1]);
    const float P2 = PLOM_UNIFORM32(pv[2]);
    const float P3 = PLOM_UNIFORM32(pv[3]);
    const float P4 = PLOM_UNIFORM32(pv[4]);
    const float P5 = PLOM_UNIFORM32(pv[5]);
    const float P6 = PLOM_UNIFORM32(pv[6]);
    const float P7 = PLOM_UNIFORM32(pv[7]);
    const float P8 = PLOM_UNIFORM32(pv[8]);
    const float P9 = PLOM_UNIFORM32(pv[9]);
    const float P10 = PLOM_UNIFORM32(pv[10]);
    const float P11 = PLOM_UNIFORM32(pv[11]);
    const float P12 = PLOM_UNIFORM32(pv[12]);
    const float P13 = PLOM_UNIFORM32(pv[13]);
    const float P14 = PLOM_UNIFORM32(pv[14]);
    const float P15 = PLOM_UNIFORM32(pv[15]);
    const float P16 = PLOM_UNIFORM32(pv[16]);
    const float P17 = PLOM_UNIFORM32(pv[17]);
    const float P18 = PLOM_UNIFORM32(pv[18]);
    const float P19 = PLOM_UNIFORM32(pv[19]);
    const float P32 = PLOM_UNIFORM32(pv[32]);
    const float P33 = PLOM_UNIFORM32(pv[33]);
    const float P34 = PLOM_UNIFORM32(pv[34]);
    const float P35 = PLOM_UNIFORM32(pv[35]);
    const float P36 = PLOM_UNIFORM32(pv[36]);
    const float P37 = PLOM_UNIFORM32(pv[37]);
    const float P38 = PLOM_UNIFORM32(pv[38]);
    const float P39 = PLOM_UNIFORM32(pv[39]);
    const float P40 = PLOM_UNIFORM32(pv[40]);
    const float P41 = PLOM_UNIFORM32(pv[41]);
    const float P42 = PLOM_UNIFORM32(pv[42]);
    const float P43 = PLOM_UNIFORM32(pv[43]);
    const float P44 = PLOM_UNIFORM32(pv[44]);
    const float P45 = PLOM_UNIFORM32(pv[45]);
    const float P46 = PLOM_UNIFORM32(pv[46]);
    const float P47 = PLOM_UNIFORM32(pv[47]);
    X[8] = 0.0f; /* reset_inc */
    X[9] = 0.0f; /* reset_inc */
    X[10] = 0.0f; /* reset_inc */
    X[11] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        const float C1 = (float)plom_cov_ld(cvk + 1);
        const float C2 = (float)plom_cov_ld(cvk + 2);
        const float C3 = (float)plom_cov_ld(cvk + 3);
        const float tf0 = tft ? tft[k * 5 + 0] : plom_sinf(plom_fdivf(((2.0f * 3.141592653589793f) * tr), 365.0f));
        float s8 = 0.0f;
        float s9 = 0.0f;
        float s10 = 0.0f;
        float s11 = 0.0f;
        const float r0_0 = plom_fdivf(((P0 * (1.0f + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0);
        const float r1_0 = P8;
        const float r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        float za0, zb0;
        plom_normal_pair32(ns, &za0, &zb0);
        const float d0_0 = sqdt * za0;
        const float d1_0 = sqdt * zb0;
        float za1, zb1;
        plom_normal_pair32(ns, &za1, &zb1);
        const float d2_0 = sqdt * za1;
        const float w0_0 = sqdt * zb1;
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - plom_fsqrtf((r0_0*X[0]))*d0_0+ plom_fsqrtf((r2_0))*d2_0 + - (r0_0*X[0])*P16*w0_0;
        const float f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + - plom_fsqrtf((r1_0*X[4]))*d1_0+ plom_fsqrtf((r0_0*X[0]))*d0_0 + + (r0_0*X[0])*P16*w0_0;
        s8 += ( + (r0_0*X[0]))*dtr + + plom_fsqrtf((r0_0*X[0]))*d0_0 + + (r0_0*X[0])*P16*w0_0;
        X[0] = plom_clipf(f0);
        const float r0_1 = plom_fdivf(((P1 * (1.0f + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1);
        const float r1_1 = P9;
        const float r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        float za2, zb2;
        plom_normal_pair32(ns, &za2, &zb2);
        const float d0_1 = sqdt * za2;
        const float d1_1 = sqdt * zb2;
        float za3, zb3;
        plom_normal_pair32(ns, &za3, &zb3);
        const float d2_1 = sqdt * za3;
        const float w0_1 = sqdt * zb3;
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - plom_fsqrtf((r0_1*X[1]))*d0_1+ plom_fsqrtf((r2_1))*d2_1 + - (r0_1*X[1])*P17*w0_1;
        const float f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + - plom_fsqrtf((r1_1*X[5]))*d1_1+ plom_fsqrtf((r0_1*X[1]))*d0_1 + + (r0_1*X[1])*P17*w0_1;
        s9 += ( + (r0_1*X[1]))*dtr + + plom_fsqrtf((r0_1*X[1]))*d0_1 + + (r0_1*X[1])*P17*w0_1;
        X[1] = plom_clipf(f1);
        const float r0_2 = plom_fdivf(((P2 * (1.0f + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2);
        const float r1_2 = P10;
        const float r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        float za4, zb4;
        plom_normal_pair32(ns, &za4, &zb4);
        const float d0_2 = sqdt * za4;
        const float d1_2 = sqdt * zb4;
        float za5, zb5;
        plom_normal_pair32(ns, &za5, &zb5);
        const float d2_2 = sqdt * za5;
        const float w0_2 = sqdt * zb5;
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - plom_fsqrtf((r0_2*X[2]))*d0_2+ plom_fsqrtf((r2_2))*d2_2 + - (r0_2*X[2])*P18*w0_2;
        const float f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + - plom_fsqrtf((r1_2*X[6]))*d1_2+ plom_fsqrtf((r0_2*X[2]))*d0_2 + + (r0_2*X[2])*P18*w0_2;
        s10 += ( + (r0_2*X[2]))*dtr + + plom_fsqrtf((r0_2*X[2]))*d0_2 + + (r0_2*X[2])*P18*w0_2;
        X[2] = plom_clipf(f2);
        const float r0_3 = plom_fdivf(((P3 * (1.0f + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3);
        const float r1_3 = P11;
        const float r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        float za6, zb6;
        plom_normal_pair32(ns, &za6, &zb6);
        const float d0_3 = sqdt * za6;
        const float d1_3 = sqdt * zb6;
        float za7, zb7;
        plom_normal_pair32(ns, &za7, &zb7);
        const float d2_3 = sqdt * za7;
        const float w0_3 = sqdt * zb7;
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - plom_fsqrtf((r0_3*X[3]))*d0_3+ plom_fsqrtf((r2_3))*d2_3 + - (r0_3*X[3])*P19*w0_3;
        const float f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + - plom_fsqrtf((r1_3*X[7]))*d1_3+ plom_fsqrtf((r0_3*X[3]))*d0_3 + + (r0_3*X[3])*P19*w0_3;
        s11 += ( + (r0_3*X[3]))*dtr + + plom_fsqrtf((r0_3*X[3]))*d0_3 + + (r0_3*X[3])*P19*w0_3;
        X[3] = plom_clipf(f3);
        X[4] = plom_clipf(f4);
        X[5] = plom_clipf(f5);
        X[6] = plom_clipf(f6);
        X[7] = plom_clipf(f7);
        const float f8 = X[8] + s8;
        X[8] = plom_clipf(f8);
        const float f9 = X[9] + s9;
        X[9] = plom_clipf(f9);
        const float f10 = X[10] + s10;
        X[10] = plom_clipf(f10);
        const float f11 = X[11] + s11;
        X[11] = plom_clipf(f11);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_no_dem_sto_f32(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const float P0 = PLOM_UNIFORM32(pv[0]);
    const float P1 = PLOM_UNIFORM32(pv[1]);
    const float P2 = PLOM_UNIFORM32(pv[2]);
    const float P3 = PLOM_UNIFORM32(pv[3]);
    const float P4 = PLOM_UNIFORM32(pv[4]);
    const float P5 = PLOM_UNIFORM32(pv[5]);
    const float P6 = PLOM_UNIFORM32(pv[6]);
    const float P7 = PLOM_UNIFORM32(pv[7]);
    const float P8 = PLOM_UNIFORM32(pv[8]);
    const float P9 = PLOM_UNIFORM32(pv[9]);
    const float P10 = PLOM_UNIFORM32(pv[10]);
    const float P11 = PLOM_UNIFORM32(pv[11]);
    const float P12 = PLOM_UNIFORM32(pv[12]);
    const float P13 = PLOM_UNIFORM32(pv[13]);
    const float P14 = PLOM_UNIFORM32(pv[14]);
    const float P15 = PLOM_UNIFORM32(pv[15]);
    const float P16 = PLOM_UNIFORM32(pv[16]);
    const float P17 = PLOM_UNIFORM32(pv[17]);
    const float P18 = PLOM_UNIFORM32(pv[18]);
    const float P19 = PLOM_UNIFORM32(pv[19]);
    const float P32 = PLOM_UNIFORM32(pv[32]);
    const float P33 = PLOM_UNIFORM32(pv[33]);
    const float P34 = PLOM_UNIFORM32(pv[34]);
    const float P35 = PLOM_UNIFORM32(pv[35]);
    const float P36 = PLOM_UNIFORM32(pv[36]);
    const float P37 = PLOM_UNIFORM32(pv[37]);
    const float P38 = PLOM_UNIFORM32(pv[38]);
    const float P39 = PLOM_UNIFORM32(pv[39]);
    const float P40 = PLOM_UNIFORM32(pv[40]);
    const float P41 = PLOM_UNIFORM32(pv[41]);
    const float P42 = PLOM_UNIFORM32(pv[42]);
    const float P43 = PLOM_UNIFORM32(pv[43]);
    const float P44 = PLOM_UNIFORM32(pv[44]);
    const float P45 = PLOM_UNIFORM32(pv[45]);
    const float P46 = PLOM_UNIFORM32(pv[46]);
    const float P47 = PLOM_UNIFORM32(pv[47]);
    X[8] = 0.0f; /* reset_inc */
    X[9] = 0.0f; /* reset_inc */
    X[10] = 0.0f; /* reset_inc */
    X[11] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        const float C1 = (float)plom_cov_ld(cvk + 1);
        const float C2 = (float)plom_cov_ld(cvk + 2);
        const float C3 = (float)plom_cov_ld(cvk + 3);
        const float tf0 = tft ? tft[k * 5 + 0] : plom_sinf(plom_fdivf(((2.0f * 3.141592653589793f) * tr), 365.0f));
        float s8 = 0.0f;
        float s9 = 0.0f;
        float s10 = 0.0f;
        float s11 = 0.0f;
        const float r0_0 = plom_fdivf(((P0 * (1.0f + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0);
        const float r1_0 = P8;
        const float r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        float za0, zb0;
        plom_normal_pair32(ns, &za0, &zb0);
        const float w0_0 = sqdt * za0;
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - (r0_0*X[0])*P16*w0_0;
        const float f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + + (r0_0*X[0])*P16*w0_0;
        s8 += ( + (r0_0*X[0]))*dtr + + (r0_0*X[0])*P16*w0_0;
        X[0] = plom_clipf(f0);
        const float r0_1 = plom_fdivf(((P1 * (1.0f + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1);
        const float r1_1 = P9;
        const float r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        const float w0_1 = sqdt * zb0;
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - (r0_1*X[1])*P17*w0_1;
        const float f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + + (r0_1*X[1])*P17*w0_1;
        s9 += ( + (r0_1*X[1]))*dtr + + (r0_1*X[1])*P17*w0_1;
        X[1] = plom_clipf(f1);
        const float r0_2 = plom_fdivf(((P2 * (1.0f + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2);
        const float r1_2 = P10;
        const float r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        float za1, zb1;
        plom_normal_pair32(ns, &za1, &zb1);
        const float w0_2 = sqdt * za1;
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - (r0_2*X[2])*P18*w0_2;
        const float f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + + (r0_2*X[2])*P18*w0_2;
        s10 += ( + (r0_2*X[2]))*dtr + + (r0_2*X[2])*P18*w0_2;
        X[2] = plom_clipf(f2);
        const float r0_3 = plom_fdivf(((P3 * (1.0f + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3);
        const float r1_3 = P11;
        const float r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        const float w0_3 = sqdt * zb1;
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - (r0_3*X[3])*P19*w0_3;
        const float f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + + (r0_3*X[3])*P19*w0_3;
        s11 += ( + (r0_3*X[3]))*dtr + + (r0_3*X[3])*P19*w0_3;
        X[3] = plom_clipf(f3);
        X[4] = plom_clipf(f4);
        X[5] = plom_clipf(f5);
        X[6] = plom_clipf(f6);
        X[7] = plom_clipf(f7);
        const float f8 = X[8] + s8;
        X[8] = plom_clipf(f8);
        const float f9 = X[9] + s9;
        X[9] = plom_clipf(f9);
        const float f10 = X[10] + s10;
        X[10] = plom_clipf(f10);
        const float f11 = X[11] + s11;
        X[11] = plom_clipf(f11);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_no_env_sto_f32(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const float P0 = PLOM_UNIFORM32(pv[0]);
    const float P1 = PLOM_UNIFORM32(pv[1]);
    const float P2 = PLOM_UNIFORM32(pv[2]);
    const float P3 = PLOM_UNIFORM32(pv[3]);
    const float P4 = PLOM_UNIFORM32(pv[4]);
    const float P5 = PLOM_UNIFORM32(pv[5]);
    const float P6 = PLOM_UNIFORM32(pv[6]);
    const float P7 = PLOM_UNIFORM32(pv[7]);
    const float P8 = PLOM_UNIFORM32(pv[8]);
    const float P9 = PLOM_UNIFORM32(pv[9]);
    const float P10 = PLOM_UNIFORM32(pv[10]);
    const float P11 = PLOM_UNIFORM32(pv[11]);
    const float P12 = PLOM_UNIFORM32(pv[12]);
    const float P13 = PLOM_UNIFORM32(pv[13]);
    const float P14 = PLOM_UNIFORM32(pv[14]);
    const float P15 = PLOM_UNIFORM32(pv[15]);
    const float P32 = PLOM_UNIFORM32(pv[32]);
    const float P33 = PLOM_UNIFORM32(pv[33]);
    const float P34 = PLOM_UNIFORM32(pv[34]);
    const float P35 = PLOM_UNIFORM32(pv[35]);
    const float P36 = PLOM_UNIFORM32(pv[36]);
    const float P37 = PLOM_UNIFORM32(pv[37]);
    const float P38 = PLOM_UNIFORM32(pv[38]);
    const float P39 = PLOM_UNIFORM32(pv[39]);
    const float P40 = PLOM_UNIFORM32(pv[40]);
    const float P41 = PLOM_UNIFORM32(pv[41]);
    const float P42 = PLOM_UNIFORM32(pv[42]);
    const float P43 = PLOM_UNIFORM32(pv[43]);
    const float P44 = PLOM_UNIFORM32(pv[44]);
    const float P45 = PLOM_UNIFORM32(pv[45]);
    const float P46 = PLOM_UNIFORM32(pv[46]);
    const float P47 = PLOM_UNIFORM32(pv[47]);
    X[8] = 0.0f; /* reset_inc */
    X[9] = 0.0f; /* reset_inc */
    X[10] = 0.0f; /* reset_inc */
    X[11] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        const float C1 = (float)plom_cov_ld(cvk + 1);
        const float C2 = (float)plom_cov_ld(cvk + 2);
        const float C3 = (float)plom_cov_ld(cvk + 3);
        const float tf0 = tft ? tft[k * 5 + 0] : plom_sinf(plom_fdivf(((2.0f * 3.141592653589793f) * tr), 365.0f));
        float s8 = 0.0f;
        float s9 = 0.0f;
        float s10 = 0.0f;
        float s11 = 0.0f;
        const float r0_0 = plom_fdivf(((P0 * (1.0f + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0);
        const float r1_0 = P8;
        const float r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        float za0, zb0;
        plom_normal_pair32(ns, &za0, &zb0);
        const float d0_0 = sqdt * za0;
        const float d1_0 = sqdt * zb0;
        float za1, zb1;
        plom_normal_pair32(ns, &za1, &zb1);
        const float d2_0 = sqdt * za1;
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - plom_fsqrtf((r0_0*X[0]))*d0_0+ plom_fsqrtf((r2_0))*d2_0;
        const float f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + - plom_fsqrtf((r1_0*X[4]))*d1_0+ plom_fsqrtf((r0_0*X[0]))*d0_0;
        s8 += ( + (r0_0*X[0]))*dtr + + plom_fsqrtf((r0_0*X[0]))*d0_0;
        X[0] = plom_clipf(f0);
        const float r0_1 = plom_fdivf(((P1 * (1.0f + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1);
        const float r1_1 = P9;
        const float r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        const float d0_1 = sqdt * zb1;
        float za2, zb2;
        plom_normal_pair32(ns, &za2, &zb2);
        const float d1_1 = sqdt * za2;
        const float d2_1 = sqdt * zb2;
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - plom_fsqrtf((r0_1*X[1]))*d0_1+ plom_fsqrtf((r2_1))*d2_1;
        const float f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + - plom_fsqrtf((r1_1*X[5]))*d1_1+ plom_fsqrtf((r0_1*X[1]))*d0_1;
        s9 += ( + (r0_1*X[1]))*dtr + + plom_fsqrtf((r0_1*X[1]))*d0_1;
        X[1] = plom_clipf(f1);
        const float r0_2 = plom_fdivf(((P2 * (1.0f + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2);
        const float r1_2 = P10;
        const float r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        float za3, zb3;
        plom_normal_pair32(ns, &za3, &zb3);
        const float d0_2 = sqdt * za3;
        const float d1_2 = sqdt * zb3;
        float za4, zb4;
        plom_normal_pair32(ns, &za4, &zb4);
        const float d2_2 = sqdt * za4;
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - plom_fsqrtf((r0_2*X[2]))*d0_2+ plom_fsqrtf((r2_2))*d2_2;
        const float f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + - plom_fsqrtf((r1_2*X[6]))*d1_2+ plom_fsqrtf((r0_2*X[2]))*d0_2;
        s10 += ( + (r0_2*X[2]))*dtr + + plom_fsqrtf((r0_2*X[2]))*d0_2;
        X[2] = plom_clipf(f2);
        const float r0_3 = plom_fdivf(((P3 * (1.0f + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3);
        const float r1_3 = P11;
        const float r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        const float d0_3 = sqdt * zb4;
        float za5, zb5;
        plom_normal_pair32(ns, &za5, &zb5);
        const float d1_3 = sqdt * za5;
        const float d2_3 = sqdt * zb5;
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - plom_fsqrtf((r0_3*X[3]))*d0_3+ plom_fsqrtf((r2_3))*d2_3;
        const float f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + - plom_fsqrtf((r1_3*X[7]))*d1_3+ plom_fsqrtf((r0_3*X[3]))*d0_3;
        s11 += ( + (r0_3*X[3]))*dtr + + plom_fsqrtf((r0_3*X[3]))*d0_3;
        X[3] = plom_clipf(f3);
        X[4] = plom_clipf(f4);
        X[5] = plom_clipf(f5);
        X[6] = plom_clipf(f6);
        X[7] = plom_clipf(f7);
        const float f8 = X[8] + s8;
        X[8] = plom_clipf(f8);
        const float f9 = X[9] + s9;
        X[9] = plom_clipf(f9);
        const float f10 = X[10] + s10;
        X[10] = plom_clipf(f10);
        const float f11 = X[11] + s11;
        X[11] = plom_clipf(f11);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_no_dem_sto_no_env_sto_f32(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft;
    const float P0 = PLOM_UNIFORM32(pv[0]);
    const float P1 = PLOM_UNIFORM32(pv[1]);
    const float P2 = PLOM_UNIFORM32(pv[2]);
    const float P3 = PLOM_UNIFORM32(pv[3]);
    const float P4 = PLOM_UNIFORM32(pv[4]);
    const float P5 = PLOM_UNIFORM32(pv[5]);
    const float P6 = PLOM_UNIFORM32(pv[6]);
    const float P7 = PLOM_UNIFORM32(pv[7]);
    const float P8 = PLOM_UNIFORM32(pv[8]);
    const float P9 = PLOM_UNIFORM32(pv[9]);
    const float P10 = PLOM_UNIFORM32(pv[10]);
    const float P11 = PLOM_UNIFORM32(pv[11]);
    const float P12 = PLOM_UNIFORM32(pv[12]);
    const float P13 = PLOM_UNIFORM32(pv[13]);
    const float P14 = PLOM_UNIFORM32(pv[14]);
    const float P15 = PLOM_UNIFORM32(pv[15]);
    const float P32 = PLOM_UNIFORM32(pv[32]);
    const float P33 = PLOM_UNIFORM32(pv[33]);
    const float P34 = PLOM_UNIFORM32(pv[34]);
    const float P35 = PLOM_UNIFORM32(pv[35]);
    const float P36 = PLOM_UNIFORM32(pv[36]);
    const float P37 = PLOM_UNIFORM32(pv[37]);
    const float P38 = PLOM_UNIFORM32(pv[38]);
    const float P39 = PLOM_UNIFORM32(pv[39]);
    const float P40 = PLOM_UNIFORM32(pv[40]);
    const float P41 = PLOM_UNIFORM32(pv[41]);
    const float P42 = PLOM_UNIFORM32(pv[42]);
    const float P43 = PLOM_UNIFORM32(pv[43]);
    const float P44 = PLOM_UNIFORM32(pv[44]);
    const float P45 = PLOM_UNIFORM32(pv[45]);
    const float P46 = PLOM_UNIFORM32(pv[46]);
    const float P47 = PLOM_UNIFORM32(pv[47]);
    X[8] = 0.0f; /* reset_inc */
    X[9] = 0.0f; /* reset_inc */
    X[10] = 0.0f; /* reset_inc */
    X[11] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        const float C1 = (float)plom_cov_ld(cvk + 1);
        const float C2 = (float)plom_cov_ld(cvk + 2);
        const float C3 = (float)plom_cov_ld(cvk + 3);
        const float tf0 = tft ? tft[k * 5 + 0] : plom_sinf(plom_fdivf(((2.0f * 3.141592653589793f) * tr), 365.0f));
        float s8 = 0.0f;
        float s9 = 0.0f;
        float s10 = 0.0f;
        float s11 = 0.0f;
        const float r0_0 = plom_fdivf(((P0 * (1.0f + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0);
        const float r1_0 = P8;
        const float r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr;
        const float f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr;
        s8 += ( + (r0_0*X[0]))*dtr;
        X[0] = plom_clipf(f0);
        const float r0_1 = plom_fdivf(((P1 * (1.0f + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1);
        const float r1_1 = P9;
        const float r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr;
        const float f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr;
        s9 += ( + (r0_1*X[1]))*dtr;
        X[1] = plom_clipf(f1);
        const float r0_2 = plom_fdivf(((P2 * (1.0f + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2);
        const float r1_2 = P10;
        const float r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr;
        const float f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr;
        s10 += ( + (r0_2*X[2]))*dtr;
        X[2] = plom_clipf(f2);
        const float r0_3 = plom_fdivf(((P3 * (1.0f + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3);
        const float r1_3 = P11;
        const float r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr;
        const float f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr;
        s11 += ( + (r0_3*X[3]))*dtr;
        X[3] = plom_clipf(f3);
        X[4] = plom_clipf(f4);
        X[5] = plom_clipf(f5);
        X[6] = plom_clipf(f6);
        X[7] = plom_clipf(f7);
        const float f8 = X[8] + s8;
        X[8] = plom_clipf(f8);
        const float f9 = X[9] + s9;
        X[9] = plom_clipf(f9);
        const float f10 = X[10] + s10;
        X[10] = plom_clipf(f10);
        const float f11 = X[11] + s11;
        X[11] = plom_clipf(f11);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_proj2obs_f32(const float *X, const double *pvd, const double *cvd, double *obs)
{
    (void)pvd; (void)cvd; (void)X;
    obs[0] = (double)X[8];
    obs[1] = (double)X[9];
    obs[2] = (double)X[10];
    obs[3] = (double)X[11];
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_psr_full_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const double P0 = lp[0];
    const double P1 = lp[1];
    const double P2 = lp[2];
    const double P3 = lp[3];
    const double P4 = lp[4];
    const double P5 = lp[5];
    const double P6 = lp[6];
    const double P7 = lp[7];
    const double P8 = lp[8];
    const double P9 = lp[9];
    const double P10 = lp[10];
    const double P11 = lp[11];
    const double P12 = lp[12];
    const double P13 = lp[13];
    const double P14 = lp[14];
    const double P15 = lp[15];
    const double P16 = lp[16];
    const double P17 = lp[17];
    const double P18 = lp[18];
    const double P19 = lp[19];
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P40 = PLOM_UNIFORM64(pv[40]);
    const double P41 = PLOM_UNIFORM64(pv[41]);
    const double P42 = PLOM_UNIFORM64(pv[42]);
    const double P43 = PLOM_UNIFORM64(pv[43]);
    const double P44 = PLOM_UNIFORM64(pv[44]);
    const double P45 = PLOM_UNIFORM64(pv[45]);
    const double P46 = PLOM_UNIFORM64(pv[46]);
    const double P47 = PLOM_UNIFORM64(pv[47]);
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double RC0 = tft ? tft[k * 5 + 1] : plom_recip_uniform(C0);
        const double RC1 = tft ? tft[k * 5 + 2] : plom_recip_uniform(C1);
        const double RC2 = tft ? tft[k * 5 + 3] : plom_recip_uniform(C2);
        const double RC3 = tft ? tft[k * 5 + 4] : plom_recip_uniform(C3);
        const double tf0 = tft ? tft[k * 5 + 0] : plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0));
        double n0_0 = 0.0;
        double n1_0 = 0.0;
        double n2_0 = 0.0;
        double n0_1 = 0.0;
        double n1_1 = 0.0;
        double n2_1 = 0.0;
        double n0_2 = 0.0;
        double n1_2 = 0.0;
        double n2_2 = 0.0;
        double n0_3 = 0.0;
        double n1_3 = 0.0;
        double n2_3 = 0.0;
        double stay0 = 0.0;
        double stay1 = 0.0;
        double stay2 = 0.0;
        double stay3 = 0.0;
        double stay4 = 0.0;
        double stay5 = 0.0;
        double stay6 = 0.0;
        double stay7 = 0.0;
        (void)n0_0; (void)n1_0; (void)n2_0; (void)n0_1; (void)n1_1; (void)n2_1; (void)n0_2; (void)n1_2; (void)n2_2; (void)n0_3; (void)n1_3; (void)n2_3;
        const double r0_0 = plom_div_rc(((P0 * (1.0 + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0, RC0);
        const double r1_0 = P8;
        const double r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        const double r0_1 = plom_div_rc(((P1 * (1.0 + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1, RC1);
        const double r1_1 = P9;
        const double r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        const double r0_2 = plom_div_rc(((P2 * (1.0 + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2, RC2);
        const double r1_2 = P10;
        const double r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        const double r0_3 = plom_div_rc(((P3 * (1.0 + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3, RC3);
        const double r1_3 = P11;
        const double r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        const double sd0_0 = P16;
        const double v0_0 = sd0_0 * sd0_0;
        const double xi0_0 = (r0_0 > 0.0) ? (v0_0 * plom_gamma_draw(ns, dt / v0_0)) / dt : 1.0;
        {
            const double sum = (r0_0 * xi0_0) * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * (((r0_0 * xi0_0) * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[0]);
            n0_0 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay0 = (p0 == p0 && cnt > 0.0) ? cnt - n0_0 : 0.0;
        }
        {
            const double sum = r1_0 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_0 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[4]);
            n1_0 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay4 = (p0 == p0 && cnt > 0.0) ? cnt - n1_0 : 0.0;
        }
        {
            const double mu = r2_0 * dt;
            n2_0 = (mu > 0.0) ? plom_poisson_draw(ns, mu) : 0.0;
        }
        const double sd0_1 = P17;
        const double v0_1 = sd0_1 * sd0_1;
        const double xi0_1 = (r0_1 > 0.0) ? (v0_1 * plom_gamma_draw(ns, dt / v0_1)) / dt : 1.0;
        {
            const double sum = (r0_1 * xi0_1) * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * (((r0_1 * xi0_1) * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[1]);
            n0_1 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay1 = (p0 == p0 && cnt > 0.0) ? cnt - n0_1 : 0.0;
        }
        {
            const double sum = r1_1 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_1 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[5]);
            n1_1 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay5 = (p0 == p0 && cnt > 0.0) ? cnt - n1_1 : 0.0;
        }
        {
            const double mu = r2_1 * dt;
            n2_1 = (mu > 0.0) ? plom_poisson_draw(ns, mu) : 0.0;
        }
        const double sd0_2 = P18;
        const double v0_2 = sd0_2 * sd0_2;
        const double xi0_2 = (r0_2 > 0.0) ? (v0_2 * plom_gamma_draw(ns, dt / v0_2)) / dt : 1.0;
        {
            const double sum = (r0_2 * xi0_2) * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * (((r0_2 * xi0_2) * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[2]);
            n0_2 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay2 = (p0 == p0 && cnt > 0.0) ? cnt - n0_2 : 0.0;
        }
        {
            const double sum = r1_2 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_2 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[6]);
            n1_2 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay6 = (p0 == p0 && cnt > 0.0) ? cnt - n1_2 : 0.0;
        }
        {
            const double mu = r2_2 * dt;
            n2_2 = (mu > 0.0) ? plom_poisson_draw(ns, mu) : 0.0;
        }
        const double sd0_3 = P19;
        const double v0_3 = sd0_3 * sd0_3;
        const double xi0_3 = (r0_3 > 0.0) ? (v0_3 * plom_gamma_draw(ns, dt / v0_3)) / dt : 1.0;
        {
            const double sum = (r0_3 * xi0_3) * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * (((r0_3 * xi0_3) * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[3]);
            n0_3 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay3 = (p0 == p0 && cnt > 0.0) ? cnt - n0_3 : 0.0;
        }
        {
            const double sum = r1_3 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_3 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[7]);
            n1_3 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay7 = (p0 == p0 && cnt > 0.0) ? cnt - n1_3 : 0.0;
        }
        {
            const double mu = r2_3 * dt;
            n2_3 = (mu > 0.0) ? plom_poisson_draw(ns, mu) : 0.0;
        }
        X[0] = stay0 + n2_0;
        X[4] = stay4 + n0_0;
        X[1] = stay1 + n2_1;
        X[5] = stay5 + n0_1;
        X[2] = stay2 + n2_2;
        X[6] = stay6 + n0_2;
        X[3] = stay3 + n2_3;
        X[7] = stay7 + n0_3;
        X[8] = X[8] + (n0_0);
        X[9] = X[9] + (n0_1);
        X[10] = X[10] + (n0_2);
        X[11] = X[11] + (n0_3);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_psr_no_env_sto_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const double P0 = lp[0];
    const double P1 = lp[1];
    const double P2 = lp[2];
    const double P3 = lp[3];
    const double P4 = lp[4];
    const double P5 = lp[5];
    const double P6 = lp[6];
    const double P7 = lp[7];
    const double P8 = lp[8];
    const double P9 = lp[9];
    const double P10 = lp[10];
    const double P11 = lp[11];
    const double P12 = lp[12];
    const double P13 = lp[13];
    const double P14 = lp[14];
    const double P15 = lp[15];
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P40 = PLOM_UNIFORM64(pv[40]);
    const double P41 = PLOM_UNIFORM64(pv[41]);
    const double P42 = PLOM_UNIFORM64(pv[42]);
    const double P43 = PLOM_UNIFORM64(pv[43]);
    const double P44 = PLOM_UNIFORM64(pv[44]);
    const double P45 = PLOM_UNIFORM64(pv[45]);
    const double P46 = PLOM_UNIFORM64(pv[46]);
    const double P47 = PLOM_UNIFORM64(pv[47]);
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double RC0 = tft ? tft[k * 5 + 1] : plom_recip_uniform(C0);
        const double RC1 = tft ? tft[k * 5 + 2] : plom_recip_uniform(C1);
        const double RC2 = tft ? tft[k * 5 + 3] : plom_recip_uniform(C2);
        const double RC3 = tft ? tft[k * 5 + 4] : plom_recip_uniform(C3);
        const double tf0 = tft ? tft[k * 5 + 0] : plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0));
        double n0_0 = 0.0;
        double n1_0 = 0.0;
        double n2_0 = 0.0;
        double n0_1 = 0.0;
        double n1_1 = 0.0;
        double n2_1 = 0.0;
        double n0_2 = 0.0;
        double n1_2 = 0.0;
        double n2_2 = 0.0;
        double n0_3 = 0.0;
        double n1_3 = 0.0;
        double n2_3 = 0.0;
        double stay0 = 0.0;
        double stay1 = 0.0;
        double stay2 = 0.0;
        double stay3 = 0.0;
        double stay4 = 0.0;
        double stay5 = 0.0;
        double stay6 = 0.0;
        double stay7 = 0.0;
        (void)n0_0; (void)n1_0; (void)n2_0; (void)n0_1; (void)n1_1; (void)n2_1; (void)n0_2; (void)n1_2; (void)n2_2; (void)n0_3; (void)n1_3; (void)n2_3;
        const double r0_0 = plom_div_rc(((P0 * (1.0 + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0, RC0);
        const double r1_0 = P8;
        const double r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        const double r0_1 = plom_div_rc(((P1 * (1.0 + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1, RC1);
        const double r1_1 = P9;
        const double r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        const double r0_2 = plom_div_rc(((P2 * (1.0 + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2, RC2);
        const double r1_2 = P10;
        const double r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        const double r0_3 = plom_div_rc(((P3 * (1.0 + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3, RC3);
        const double r1_3 = P11;
        const double r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        const double xi0_0 = 1.0;
        {
            const double sum = (r0_0 * xi0_0) * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * (((r0_0 * xi0_0) * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[0]);
            n0_0 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay0 = (p0 == p0 && cnt > 0.0) ? cnt - n0_0 : 0.0;
        }
        {
            const double sum = r1_0 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_0 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[4]);
            n1_0 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay4 = (p0 == p0 && cnt > 0.0) ? cnt - n1_0 : 0.0;
        }
        {
            const double mu = r2_0 * dt;
            n2_0 = (mu > 0.0) ? plom_poisson_draw(ns, mu) : 0.0;
        }
        const double xi0_1 = 1.0;
        {
            const double sum = (r0_1 * xi0_1) * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * (((r0_1 * xi0_1) * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[1]);
            n0_1 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay1 = (p0 == p0 && cnt > 0.0) ? cnt - n0_1 : 0.0;
        }
        {
            const double sum = r1_1 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_1 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[5]);
            n1_1 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay5 = (p0 == p0 && cnt > 0.0) ? cnt - n1_1 : 0.0;
        }
        {
            const double mu = r2_1 * dt;
            n2_1 = (mu > 0.0) ? plom_poisson_draw(ns, mu) : 0.0;
        }
        const double xi0_2 = 1.0;
        {
            const double sum = (r0_2 * xi0_2) * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * (((r0_2 * xi0_2) * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[2]);
            n0_2 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay2 = (p0 == p0 && cnt > 0.0) ? cnt - n0_2 : 0.0;
        }
        {
            const double sum = r1_2 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_2 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[6]);
            n1_2 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay6 = (p0 == p0 && cnt > 0.0) ? cnt - n1_2 : 0.0;
        }
        {
            const double mu = r2_2 * dt;
            n2_2 = (mu > 0.0) ? plom_poisson_draw(ns, mu) : 0.0;
        }
        const double xi0_3 = 1.0;
        {
            const double sum = (r0_3 * xi0_3) * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * (((r0_3 * xi0_3) * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[3]);
            n0_3 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay3 = (p0 == p0 && cnt > 0.0) ? cnt - n0_3 : 0.0;
        }
        {
            const double sum = r1_3 * dt;
            const double p0 = (sum > 0.0) ? (1.0 - plom_exp(-sum)) * ((r1_3 * dt) / sum) : 0.0;
            const double cnt = plom_floor(X[7]);
            n1_3 = (cnt > 0.0 && p0 > 0.0) ? plom_binom_draw(ns, cnt, p0) : 0.0;
            stay7 = (p0 == p0 && cnt > 0.0) ? cnt - n1_3 : 0.0;
        }
        {
            const double mu = r2_3 * dt;
            n2_3 = (mu > 0.0) ? plom_poisson_draw(ns, mu) : 0.0;
        }
        X[0] = stay0 + n2_0;
        X[4] = stay4 + n0_0;
        X[1] = stay1 + n2_1;
        X[5] = stay5 + n0_1;
        X[2] = stay2 + n2_2;
        X[6] = stay6 + n0_2;
        X[3] = stay3 + n2_3;
        X[7] = stay7 + n0_3;
        X[8] = X[8] + (n0_0);
        X[9] = X[9] + (n0_1);
        X[10] = X[10] + (n0_2);
        X[11] = X[11] + (n0_3);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_ode_rhs_pp(double t, const double *X, const double *pv, const double *lp, const double *cv, const plom_spline_set *sp, double *f)
{
    (void)pv; (void)lp; (void)cv; (void)sp;
    const double tr = t; (void)tr;
    const double P0 = lp[0];
    const double P1 = lp[1];
    const double P2 = lp[2];
    const double P3 = lp[3];
    const double P4 = lp[4];
    const double P5 = lp[5];
    const double P6 = lp[6];
    const double P7 = lp[7];
    const double P8 = lp[8];
    const double P9 = lp[9];
    const double P10 = lp[10];
    const double P11 = lp[11];
    const double P12 = lp[12];
    const double P13 = lp[13];
    const double P14 = lp[14];
    const double P15 = lp[15];
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P40 = PLOM_UNIFORM64(pv[40]);
    const double P41 = PLOM_UNIFORM64(pv[41]);
    const double P42 = PLOM_UNIFORM64(pv[42]);
    const double P43 = PLOM_UNIFORM64(pv[43]);
    const double P44 = PLOM_UNIFORM64(pv[44]);
    const double P45 = PLOM_UNIFORM64(pv[45]);
    const double P46 = PLOM_UNIFORM64(pv[46]);
    const double P47 = PLOM_UNIFORM64(pv[47]);
    const double C0 = sp ? plom_spline_at(sp, 0, t) : plom_cov_ld(cv + 0);
    const double C1 = sp ? plom_spline_at(sp, 1, t) : plom_cov_ld(cv + 1);
    const double C2 = sp ? plom_spline_at(sp, 2, t) : plom_cov_ld(cv + 2);
    const double C3 = sp ? plom_spline_at(sp, 3, t) : plom_cov_ld(cv + 3);
    const double tf0 = plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0));
    const double r0_0 = (((P0 * (1.0 + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])) / C0);
    const double r1_0 = P8;
    const double r2_0 = (P12 * ((C0 - X[0]) - X[4]));
    const double r0_1 = (((P1 * (1.0 + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])) / C1);
    const double r1_1 = P9;
    const double r2_1 = (P13 * ((C1 - X[1]) - X[5]));
    const double r0_2 = (((P2 * (1.0 + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])) / C2);
    const double r1_2 = P10;
    const double r2_2 = (P14 * ((C2 - X[2]) - X[6]));
    const double r0_3 = (((P3 * (1.0 + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])) / C3);
    const double r1_3 = P11;
    const double r2_3 = (P15 * ((C3 - X[3]) - X[7]));
    f[0] =  - (r0_0*X[0]) + (r2_0);
    f[4] =  - (r1_0*X[4]) + (r0_0*X[0]);
    f[1] =  - (r0_1*X[1]) + (r2_1);
    f[5] =  - (r1_1*X[5]) + (r0_1*X[1]);
    f[2] =  - (r0_2*X[2]) + (r2_2);
    f[6] =  - (r1_2*X[6]) + (r0_2*X[2]);
    f[3] =  - (r0_3*X[3]) + (r2_3);
    f[7] =  - (r1_3*X[7]) + (r0_3*X[3]);
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r0_0*X[0]);
        f[8] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r0_1*X[1]);
        f[9] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r0_2*X[2]);
        f[10] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r0_3*X[3]);
        f[11] = sum_inc;
    }
}

#define PLOM_ODE_FN plom_sirs4_ode_integrate_pp
#define PLOM_ODE_RHS plom_sirs4_ode_rhs_pp
#define PLOM_ODE_N 12
#define PLOM_ODE_QUAL static __device__ inline __attribute__((always_inline))
#include "plom_ode_impl.h"

static __device__ inline __attribute__((always_inline)) int plom_sirs4_interval_ode_pp_f64(double *X, const double *pv, const double *lp, const double *cv, const plom_spline_set *sp, double t0, double t1, double dt, double eps_abs, double eps_rel)
{
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    return plom_sirs4_ode_integrate_pp(X, pv, lp, cv, sp, t0, t1, dt, eps_abs, eps_rel);
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_full_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const double P0 = lp[0];
    const double P1 = lp[1];
    const double P2 = lp[2];
    const double P3 = lp[3];
    const double P4 = lp[4];
    const double P5 = lp[5];
    const double P6 = lp[6];
    const double P7 = lp[7];
    const double P8 = lp[8];
    const double P9 = lp[9];
    const double P10 = lp[10];
    const double P11 = lp[11];
    const double P12 = lp[12];
    const double P13 = lp[13];
    const double P14 = lp[14];
    const double P15 = lp[15];
    const double P16 = lp[16];
    const double P17 = lp[17];
    const double P18 = lp[18];
    const double P19 = lp[19];
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P40 = PLOM_UNIFORM64(pv[40]);
    const double P41 = PLOM_UNIFORM64(pv[41]);
    const double P42 = PLOM_UNIFORM64(pv[42]);
    const double P43 = PLOM_UNIFORM64(pv[43]);
    const double P44 = PLOM_UNIFORM64(pv[44]);
    const double P45 = PLOM_UNIFORM64(pv[45]);
    const double P46 = PLOM_UNIFORM64(pv[46]);
    const double P47 = PLOM_UNIFORM64(pv[47]);
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double RC0 = tft ? tft[k * 5 + 1] : plom_recip_uniform(C0);
        const double RC1 = tft ? tft[k * 5 + 2] : plom_recip_uniform(C1);
        const double RC2 = tft ? tft[k * 5 + 3] : plom_recip_uniform(C2);
        const double RC3 = tft ? tft[k * 5 + 4] : plom_recip_uniform(C3);
        const double tf0 = tft ? tft[k * 5 + 0] : plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0));
        double s8 = 0.0;
        double s9 = 0.0;
        double s10 = 0.0;
        double s11 = 0.0;
        const double r0_0 = plom_div_rc(((P0 * (1.0 + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0, RC0);
        const double r1_0 = P8;
        const double r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        double za0, zb0;
        plom_normal_pair64(ns, &za0, &zb0);
        const double d0_0 = sqdt * za0;
        const double d1_0 = sqdt * zb0;
        double za1, zb1;
        plom_normal_pair64(ns, &za1, &zb1);
        const double d2_0 = sqdt * za1;
        const double w0_0 = sqdt * zb1;
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - plom_sqrt_rate((r0_0*X[0]))*d0_0+ plom_sqrt_rate((r2_0))*d2_0 + - (r0_0*X[0])*P16*w0_0;
        const double f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + - plom_sqrt_rate((r1_0*X[4]))*d1_0+ plom_sqrt_rate((r0_0*X[0]))*d0_0 + + (r0_0*X[0])*P16*w0_0;
        s8 += ( + (r0_0*X[0]))*dtr + + plom_sqrt_rate((r0_0*X[0]))*d0_0 + + (r0_0*X[0])*P16*w0_0;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        const double r0_1 = plom_div_rc(((P1 * (1.0 + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1, RC1);
        const double r1_1 = P9;
        const double r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        double za2, zb2;
        plom_normal_pair64(ns, &za2, &zb2);
        const double d0_1 = sqdt * za2;
        const double d1_1 = sqdt * zb2;
        double za3, zb3;
        plom_normal_pair64(ns, &za3, &zb3);
        const double d2_1 = sqdt * za3;
        const double w0_1 = sqdt * zb3;
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - plom_sqrt_rate((r0_1*X[1]))*d0_1+ plom_sqrt_rate((r2_1))*d2_1 + - (r0_1*X[1])*P17*w0_1;
        const double f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + - plom_sqrt_rate((r1_1*X[5]))*d1_1+ plom_sqrt_rate((r0_1*X[1]))*d0_1 + + (r0_1*X[1])*P17*w0_1;
        s9 += ( + (r0_1*X[1]))*dtr + + plom_sqrt_rate((r0_1*X[1]))*d0_1 + + (r0_1*X[1])*P17*w0_1;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        const double r0_2 = plom_div_rc(((P2 * (1.0 + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2, RC2);
        const double r1_2 = P10;
        const double r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        double za4, zb4;
        plom_normal_pair64(ns, &za4, &zb4);
        const double d0_2 = sqdt * za4;
        const double d1_2 = sqdt * zb4;
        double za5, zb5;
        plom_normal_pair64(ns, &za5, &zb5);
        const double d2_2 = sqdt * za5;
        const double w0_2 = sqdt * zb5;
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - plom_sqrt_rate((r0_2*X[2]))*d0_2+ plom_sqrt_rate((r2_2))*d2_2 + - (r0_2*X[2])*P18*w0_2;
        const double f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + - plom_sqrt_rate((r1_2*X[6]))*d1_2+ plom_sqrt_rate((r0_2*X[2]))*d0_2 + + (r0_2*X[2])*P18*w0_2;
        s10 += ( + (r0_2*X[2]))*dtr + + plom_sqrt_rate((r0_2*X[2]))*d0_2 + + (r0_2*X[2])*P18*w0_2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        const double r0_3 = plom_div_rc(((P3 * (1.0 + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3, RC3);
        const double r1_3 = P11;
        const double r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        double za6, zb6;
        plom_normal_pair64(ns, &za6, &zb6);
        const double d0_3 = sqdt * za6;
        const double d1_3 = sqdt * zb6;
        double za7, zb7;
        plom_normal_pair64(ns, &za7, &zb7);
        const double d2_3 = sqdt * za7;
        const double w0_3 = sqdt * zb7;
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - plom_sqrt_rate((r0_3*X[3]))*d0_3+ plom_sqrt_rate((r2_3))*d2_3 + - (r0_3*X[3])*P19*w0_3;
        const double f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + - plom_sqrt_rate((r1_3*X[7]))*d1_3+ plom_sqrt_rate((r0_3*X[3]))*d0_3 + + (r0_3*X[3])*P19*w0_3;
        s11 += ( + (r0_3*X[3]))*dtr + + plom_sqrt_rate((r0_3*X[3]))*d0_3 + + (r0_3*X[3])*P19*w0_3;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        const double f8 = X[8] + s8;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        const double f9 = X[9] + s9;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        const double f10 = X[10] + s10;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        const double f11 = X[11] + s11;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_no_dem_sto_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const double P0 = lp[0];
    const double P1 = lp[1];
    const double P2 = lp[2];
    const double P3 = lp[3];
    const double P4 = lp[4];
    const double P5 = lp[5];
    const double P6 = lp[6];
    const double P7 = lp[7];
    const double P8 = lp[8];
    const double P9 = lp[9];
    const double P10 = lp[10];
    const double P11 = lp[11];
    const double P12 = lp[12];
    const double P13 = lp[13];
    const double P14 = lp[14];
    const double P15 = lp[15];
    const double P16 = lp[16];
    const double P17 = lp[17];
    const double P18 = lp[18];
    const double P19 = lp[19];
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P40 = PLOM_UNIFORM64(pv[40]);
    const double P41 = PLOM_UNIFORM64(pv[41]);
    const double P42 = PLOM_UNIFORM64(pv[42]);
    const double P43 = PLOM_UNIFORM64(pv[43]);
    const double P44 = PLOM_UNIFORM64(pv[44]);
    const double P45 = PLOM_UNIFORM64(pv[45]);
    const double P46 = PLOM_UNIFORM64(pv[46]);
    const double P47 = PLOM_UNIFORM64(pv[47]);
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double RC0 = tft ? tft[k * 5 + 1] : plom_recip_uniform(C0);
        const double RC1 = tft ? tft[k * 5 + 2] : plom_recip_uniform(C1);
        const double RC2 = tft ? tft[k * 5 + 3] : plom_recip_uniform(C2);
        const double RC3 = tft ? tft[k * 5 + 4] : plom_recip_uniform(C3);
        const double tf0 = tft ? tft[k * 5 + 0] : plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0));
        double s8 = 0.0;
        double s9 = 0.0;
        double s10 = 0.0;
        double s11 = 0.0;
        const double r0_0 = plom_div_rc(((P0 * (1.0 + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0, RC0);
        const double r1_0 = P8;
        const double r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        double za0, zb0;
        plom_normal_pair64(ns, &za0, &zb0);
        const double w0_0 = sqdt * za0;
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - (r0_0*X[0])*P16*w0_0;
        const double f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + + (r0_0*X[0])*P16*w0_0;
        s8 += ( + (r0_0*X[0]))*dtr + + (r0_0*X[0])*P16*w0_0;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        const double r0_1 = plom_div_rc(((P1 * (1.0 + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1, RC1);
        const double r1_1 = P9;
        const double r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        const double w0_1 = sqdt * zb0;
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - (r0_1*X[1])*P17*w0_1;
        const double f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + + (r0_1*X[1])*P17*w0_1;
        s9 += ( + (r0_1*X[1]))*dtr + + (r0_1*X[1])*P17*w0_1;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        const double r0_2 = plom_div_rc(((P2 * (1.0 + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2, RC2);
        const double r1_2 = P10;
        const double r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        double za1, zb1;
        plom_normal_pair64(ns, &za1, &zb1);
        const double w0_2 = sqdt * za1;
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - (r0_2*X[2])*P18*w0_2;
        const double f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + + (r0_2*X[2])*P18*w0_2;
        s10 += ( + (r0_2*X[2]))*dtr + + (r0_2*X[2])*P18*w0_2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        const double r0_3 = plom_div_rc(((P3 * (1.0 + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3, RC3);
        const double r1_3 = P11;
        const double r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        const double w0_3 = sqdt * zb1;
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - (r0_3*X[3])*P19*w0_3;
        const double f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + + (r0_3*X[3])*P19*w0_3;
        s11 += ( + (r0_3*X[3]))*dtr + + (r0_3*X[3])*P19*w0_3;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        const double f8 = X[8] + s8;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        const double f9 = X[9] + s9;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        const double f10 = X[10] + s10;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        const double f11 = X[11] + s11;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_no_env_sto_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const double P0 = lp[0];
    const double P1 = lp[1];
    const double P2 = lp[2];
    const double P3 = lp[3];
    const double P4 = lp[4];
    const double P5 = lp[5];
    const double P6 = lp[6];
    const double P7 = lp[7];
    const double P8 = lp[8];
    const double P9 = lp[9];
    const double P10 = lp[10];
    const double P11 = lp[11];
    const double P12 = lp[12];
    const double P13 = lp[13];
    const double P14 = lp[14];
    const double P15 = lp[15];
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P40 = PLOM_UNIFORM64(pv[40]);
    const double P41 = PLOM_UNIFORM64(pv[41]);
    const double P42 = PLOM_UNIFORM64(pv[42]);
    const double P43 = PLOM_UNIFORM64(pv[43]);
    const double P44 = PLOM_UNIFORM64(pv[44]);
    const double P45 = PLOM_UNIFORM64(pv[45]);
    const double P46 = PLOM_UNIFORM64(pv[46]);
    const double P47 = PLOM_UNIFORM64(pv[47]);
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double RC0 = tft ? tft[k * 5 + 1] : plom_recip_uniform(C0);
        const double RC1 = tft ? tft[k * 5 + 2] : plom_recip_uniform(C1);
        const double RC2 = tft ? tft[k * 5 + 3] : plom_recip_uniform(C2);
        const double RC3 = tft ? tft[k * 5 + 4] : plom_recip_uniform(C3);
        const double tf0 = tft ? tft[k * 5 + 0] : plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0));
        double s8 = 0.0;
        double s9 = 0.0;
        double s10 = 0.0;
        double s11 = 0.0;
        const double r0_0 = plom_div_rc(((P0 * (1.0 + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0, RC0);
        const double r1_0 = P8;
        const double r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        double za0, zb0;
        plom_normal_pair64(ns, &za0, &zb0);
        const double d0_0 = sqdt * za0;
        const double d1_0 = sqdt * zb0;
        double za1, zb1;
        plom_normal_pair64(ns, &za1, &zb1);
        const double d2_0 = sqdt * za1;
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - plom_sqrt_rate((r0_0*X[0]))*d0_0+ plom_sqrt_rate((r2_0))*d2_0;
        const double f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + - plom_sqrt_rate((r1_0*X[4]))*d1_0+ plom_sqrt_rate((r0_0*X[0]))*d0_0;
        s8 += ( + (r0_0*X[0]))*dtr + + plom_sqrt_rate((r0_0*X[0]))*d0_0;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        const double r0_1 = plom_div_rc(((P1 * (1.0 + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1, RC1);
        const double r1_1 = P9;
        const double r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        const double d0_1 = sqdt * zb1;
        double za2, zb2;
        plom_normal_pair64(ns, &za2, &zb2);
        const double d1_1 = sqdt * za2;
        const double d2_1 = sqdt * zb2;
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - plom_sqrt_rate((r0_1*X[1]))*d0_1+ plom_sqrt_rate((r2_1))*d2_1;
        const double f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + - plom_sqrt_rate((r1_1*X[5]))*d1_1+ plom_sqrt_rate((r0_1*X[1]))*d0_1;
        s9 += ( + (r0_1*X[1]))*dtr + + plom_sqrt_rate((r0_1*X[1]))*d0_1;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        const double r0_2 = plom_div_rc(((P2 * (1.0 + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2, RC2);
        const double r1_2 = P10;
        const double r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        double za3, zb3;
        plom_normal_pair64(ns, &za3, &zb3);
        const double d0_2 = sqdt * za3;
        const double d1_2 = sqdt * zb3;
        double za4, zb4;
        plom_normal_pair64(ns, &za4, &zb4);
        const double d2_2 = sqdt * za4;
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - plom_sqrt_rate((r0_2*X[2]))*d0_2+ plom_sqrt_rate((r2_2))*d2_2;
        const double f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + - plom_sqrt_rate((r1_2*X[6]))*d1_2+ plom_sqrt_rate((r0_2*X[2]))*d0_2;
        s10 += ( + (r0_2*X[2]))*dtr + + plom_sqrt_rate((r0_2*X[2]))*d0_2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        const double r0_3 = plom_div_rc(((P3 * (1.0 + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3, RC3);
        const double r1_3 = P11;
        const double r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        const double d0_3 = sqdt * zb4;
        double za5, zb5;
        plom_normal_pair64(ns, &za5, &zb5);
        const double d1_3 = sqdt * za5;
        const double d2_3 = sqdt * zb5;
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - plom_sqrt_rate((r0_3*X[3]))*d0_3+ plom_sqrt_rate((r2_3))*d2_3;
        const double f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + - plom_sqrt_rate((r1_3*X[7]))*d1_3+ plom_sqrt_rate((r0_3*X[3]))*d0_3;
        s11 += ( + (r0_3*X[3]))*dtr + + plom_sqrt_rate((r0_3*X[3]))*d0_3;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        const double f8 = X[8] + s8;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        const double f9 = X[9] + s9;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        const double f10 = X[10] + s10;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        const double f11 = X[11] + s11;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_no_dem_sto_no_env_sto_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
{
    const double dtr = (double)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const double P0 = lp[0];
    const double P1 = lp[1];
    const double P2 = lp[2];
    const double P3 = lp[3];
    const double P4 = lp[4];
    const double P5 = lp[5];
    const double P6 = lp[6];
    const double P7 = lp[7];
    const double P8 = lp[8];
    const double P9 = lp[9];
    const double P10 = lp[10];
    const double P11 = lp[11];
    const double P12 = lp[12];
    const double P13 = lp[13];
    const double P14 = lp[14];
    const double P15 = lp[15];
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P40 = PLOM_UNIFORM64(pv[40]);
    const double P41 = PLOM_UNIFORM64(pv[41]);
    const double P42 = PLOM_UNIFORM64(pv[42]);
    const double P43 = PLOM_UNIFORM64(pv[43]);
    const double P44 = PLOM_UNIFORM64(pv[44]);
    const double P45 = PLOM_UNIFORM64(pv[45]);
    const double P46 = PLOM_UNIFORM64(pv[46]);
    const double P47 = PLOM_UNIFORM64(pv[47]);
    X[8] = 0.0; /* reset_inc */
    X[9] = 0.0; /* reset_inc */
    X[10] = 0.0; /* reset_inc */
    X[11] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double RC0 = tft ? tft[k * 5 + 1] : plom_recip_uniform(C0);
        const double RC1 = tft ? tft[k * 5 + 2] : plom_recip_uniform(C1);
        const double RC2 = tft ? tft[k * 5 + 3] : plom_recip_uniform(C2);
        const double RC3 = tft ? tft[k * 5 + 4] : plom_recip_uniform(C3);
        const double tf0 = tft ? tft[k * 5 + 0] : plom_sin((((2.0 * 3.141592653589793) * tr) / 365.0));
        double s8 = 0.0;
        double s9 = 0.0;
        double s10 = 0.0;
        double s11 = 0.0;
        const double r0_0 = plom_div_rc(((P0 * (1.0 + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0, RC0);
        const double r1_0 = P8;
        const double r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr;
        const double f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr;
        s8 += ( + (r0_0*X[0]))*dtr;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        const double r0_1 = plom_div_rc(((P1 * (1.0 + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1, RC1);
        const double r1_1 = P9;
        const double r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr;
        const double f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr;
        s9 += ( + (r0_1*X[1]))*dtr;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        const double r0_2 = plom_div_rc(((P2 * (1.0 + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2, RC2);
        const double r1_2 = P10;
        const double r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr;
        const double f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr;
        s10 += ( + (r0_2*X[2]))*dtr;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        const double r0_3 = plom_div_rc(((P3 * (1.0 + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3, RC3);
        const double r1_3 = P11;
        const double r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr;
        const double f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr;
        s11 += ( + (r0_3*X[3]))*dtr;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        const double f8 = X[8] + s8;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        const double f9 = X[9] + s9;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        const double f10 = X[10] + s10;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        const double f11 = X[11] + s11;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_proj2obs_pp_f64(const double *X, const double *pvd, const double *lpd, const double *cvd, double *obs)
{
    (void)pvd; (void)lpd; (void)cvd; (void)X;
    obs[0] = (double)X[8];
    obs[1] = (double)X[9];
    obs[2] = (double)X[10];
    obs[3] = (double)X[11];
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_full_pp_f32(float *X, const float *pv, const float *lp, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const float P0 = lp[0];
    const float P1 = lp[1];
    const float P2 = lp[2];
    const float P3 = lp[3];
    const float P4 = lp[4];
    const float P5 = lp[5];
    const float P6 = lp[6];
    const float P7 = lp[7];
    const float P8 = lp[8];
    const float P9 = lp[9];
    const float P10 = lp[10];
    const float P11 = lp[11];
    const float P12 = lp[12];
    const float P13 = lp[13];
    const float P14 = lp[14];
    const float P15 = lp[15];
    const float P16 = lp[16];
    const float P17 = lp[17];
    const float P18 = lp[18];
    const float P19 = lp[19];
    const float P32 = PLOM_UNIFORM32(pv[32]);
    const float P33 = PLOM_UNIFORM32(pv[33]);
    const float P34 = PLOM_UNIFORM32(pv[34]);
    const float P35 = PLOM_UNIFORM32(pv[35]);
    const float P36 = PLOM_UNIFORM32(pv[36]);
    const float P37 = PLOM_UNIFORM32(pv[37]);
    const float P38 = PLOM_UNIFORM32(pv[38]);
    const float P39 = PLOM_UNIFORM32(pv[39]);
    const float P40 = PLOM_UNIFORM32(pv[40]);
    const float P41 = PLOM_UNIFORM32(pv[41]);
    const float P42 = PLOM_UNIFORM32(pv[42]);
    const float P43 = PLOM_UNIFORM32(pv[43]);
    const float P44 = PLOM_UNIFORM32(pv[44]);
    const float P45 = PLOM_UNIFORM32(pv[45]);
    const float P46 = PLOM_UNIFORM32(pv[46]);
    const float P47 = PLOM_UNIFORM32(pv[47]);
    X[8] = 0.0f; /* reset_inc */
    X[9] = 0.0f; /* reset_inc */
    X[10] = 0.0f; /* reset_inc */
    X[11] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        const float C1 = (float)plom_cov_ld(cvk + 1);
        const float C2 = (float)plom_cov_ld(cvk + 2);
        const float C3 = (float)plom_cov_ld(cvk + 3);
        const float tf0 = tft ? tft[k * 5 + 0] : plom_sinf(plom_fdivf(((2.0f * 3.141592653589793f) * tr), 365.0f));
        float s8 = 0.0f;
        float s9 = 0.0f;
        float s10 = 0.0f;
        float s11 = 0.0f;
        const float r0_0 = plom_fdivf(((P0 * (1.0f + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0);
        const float r1_0 = P8;
        const float r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        float za0, zb0;
        plom_normal_pair32(ns, &za0, &zb0);
        const float d0_0 = sqdt * za0;
        const float d1_0 = sqdt * zb0;
        float za1, zb1;
        plom_normal_pair32(ns, &za1, &zb1);
        const float d2_0 = sqdt * za1;
        const float w0_0 = sqdt * zb1;
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - plom_fsqrtf((r0_0*X[0]))*d0_0+ plom_fsqrtf((r2_0))*d2_0 + - (r0_0*X[0])*P16*w0_0;
        const float f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + - plom_fsqrtf((r1_0*X[4]))*d1_0+ plom_fsqrtf((r0_0*X[0]))*d0_0 + + (r0_0*X[0])*P16*w0_0;
        s8 += ( + (r0_0*X[0]))*dtr + + plom_fsqrtf((r0_0*X[0]))*d0_0 + + (r0_0*X[0])*P16*w0_0;
        X[0] = plom_clipf(f0);
        const float r0_1 = plom_fdivf(((P1 * (1.0f + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1);
        const float r1_1 = P9;
        const float r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        float za2, zb2;
        plom_normal_pair32(ns, &za2, &zb2);
        const float d0_1 = sqdt * za2;
        const float d1_1 = sqdt * zb2;
        float za3, zb3;
        plom_normal_pair32(ns, &za3, &zb3);
        const float d2_1 = sqdt * za3;
        const float w0_1 = sqdt * zb3;
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - plom_fsqrtf((r0_1*X[1]))*d0_1+ plom_fsqrtf((r2_1))*d2_1 + - (r0_1*X[1])*P17*w0_1;
        const float f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + - plom_fsqrtf((r1_1*X[5]))*d1_1+ plom_fsqrtf((r0_1*X[1]))*d0_1 + + (r0_1*X[1])*P17*w0_1;
        s9 += ( + (r0_1*X[1]))*dtr + + plom_fsqrtf((r0_1*X[1]))*d0_1 + + (r0_1*X[1])*P17*w0_1;
        X[1] = plom_clipf(f1);
        const float r0_2 = plom_fdivf(((P2 * (1.0f + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2);
        const float r1_2 = P10;
        const float r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        float za4, zb4;
        plom_normal_pair32(ns, &za4, &zb4);
        const float d0_2 = sqdt * za4;
        const float d1_2 = sqdt * zb4;
        float za5, zb5;
        plom_normal_pair32(ns, &za5, &zb5);
        const float d2_2 = sqdt * za5;
        const float w0_2 = sqdt * zb5;
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - plom_fsqrtf((r0_2*X[2]))*d0_2+ plom_fsqrtf((r2_2))*d2_2 + - (r0_2*X[2])*P18*w0_2;
        const float f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + - plom_fsqrtf((r1_2*X[6]))*d1_2+ plom_fsqrtf((r0_2*X[2]))*d0_2 + + (r0_2*X[2])*P18*w0_2;
        s10 += ( + (r0_2*X[2]))*dtr + + plom_fsqrtf((r0_2*X[2]))*d0_2 + + (r0_2*X[2])*P18*w0_2;
        X[2] = plom_clipf(f2);
        const float r0_3 = plom_fdivf(((P3 * (1.0f + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3);
        const float r1_3 = P11;
        const float r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        float za6, zb6;
        plom_normal_pair32(ns, &za6, &zb6);
        const float d0_3 = sqdt * za6;
        const float d1_3 = sqdt * zb6;
        float za7, zb7;
        plom_normal_pair32(ns, &za7, &zb7);
        const float d2_3 = sqdt * za7;
        const float w0_3 = sqdt * zb7;
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - plom_fsqrtf((r0_3*X[3]))*d0_3+ plom_fsqrtf((r2_3))*d2_3 + - (r0_3*X[3])*P19*w0_3;
        const float f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + - plom_fsqrtf((r1_3*X[7]))*d1_3+ plom_fsqrtf((r0_3*X[3]))*d0_3 + + (r0_3*X[3])*P19*w0_3;
        s11 += ( + (r0_3*X[3]))*dtr + + plom_fsqrtf((r0_3*X[3]))*d0_3 + + (r0_3*X[3])*P19*w0_3;
        X[3] = plom_clipf(f3);
        X[4] = plom_clipf(f4);
        X[5] = plom_clipf(f5);
        X[6] = plom_clipf(f6);
        X[7] = plom_clipf(f7);
        const float f8 = X[8] + s8;
        X[8] = plom_clipf(f8);
        const float f9 = X[9] + s9;
        X[9] = plom_clipf(f9);
        const float f10 = X[10] + s10;
        X[10] = plom_clipf(f10);
        const float f11 = X[11] + s11;
        X[11] = plom_clipf(f11);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_no_dem_sto_pp_f32(float *X, const float *pv, const float *lp, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const float P0 = lp[0];
    const float P1 = lp[1];
    const float P2 = lp[2];
    const float P3 = lp[3];
    const float P4 = lp[4];
    const float P5 = lp[5];
    const float P6 = lp[6];
    const float P7 = lp[7];
    const float P8 = lp[8];
    const float P9 = lp[9];
    const float P10 = lp[10];
    const float P11 = lp[11];
    const float P12 = lp[12];
    const float P13 = lp[13];
    const float P14 = lp[14];
    const float P15 = lp[15];
    const float P16 = lp[16];
    const float P17 = lp[17];
    const float P18 = lp[18];
    const float P19 = lp[19];
    const float P32 = PLOM_UNIFORM32(pv[32]);
    const float P33 = PLOM_UNIFORM32(pv[33]);
    const float P34 = PLOM_UNIFORM32(pv[34]);
    const float P35 = PLOM_UNIFORM32(pv[35]);
    const float P36 = PLOM_UNIFORM32(pv[36]);
    const float P37 = PLOM_UNIFORM32(pv[37]);
    const float P38 = PLOM_UNIFORM32(pv[38]);
    const float P39 = PLOM_UNIFORM32(pv[39]);
    const float P40 = PLOM_UNIFORM32(pv[40]);
    const float P41 = PLOM_UNIFORM32(pv[41]);
    const float P42 = PLOM_UNIFORM32(pv[42]);
    const float P43 = PLOM_UNIFORM32(pv[43]);
    const float P44 = PLOM_UNIFORM32(pv[44]);
    const float P45 = PLOM_UNIFORM32(pv[45]);
    const float P46 = PLOM_UNIFORM32(pv[46]);
    const float P47 = PLOM_UNIFORM32(pv[47]);
    X[8] = 0.0f; /* reset_inc */
    X[9] = 0.0f; /* reset_inc */
    X[10] = 0.0f; /* reset_inc */
    X[11] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        const float C1 = (float)plom_cov_ld(cvk + 1);
        const float C2 = (float)plom_cov_ld(cvk + 2);
        const float C3 = (float)plom_cov_ld(cvk + 3);
        const float tf0 = tft ? tft[k * 5 + 0] : plom_sinf(plom_fdivf(((2.0f * 3.141592653589793f) * tr), 365.0f));
        float s8 = 0.0f;
        float s9 = 0.0f;
        float s10 = 0.0f;
        float s11 = 0.0f;
        const float r0_0 = plom_fdivf(((P0 * (1.0f + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0);
        const float r1_0 = P8;
        const float r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        float za0, zb0;
        plom_normal_pair32(ns, &za0, &zb0);
        const float w0_0 = sqdt * za0;
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - (r0_0*X[0])*P16*w0_0;
        const float f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + + (r0_0*X[0])*P16*w0_0;
        s8 += ( + (r0_0*X[0]))*dtr + + (r0_0*X[0])*P16*w0_0;
        X[0] = plom_clipf(f0);
        const float r0_1 = plom_fdivf(((P1 * (1.0f + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1);
        const float r1_1 = P9;
        const float r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        const float w0_1 = sqdt * zb0;
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - (r0_1*X[1])*P17*w0_1;
        const float f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + + (r0_1*X[1])*P17*w0_1;
        s9 += ( + (r0_1*X[1]))*dtr + + (r0_1*X[1])*P17*w0_1;
        X[1] = plom_clipf(f1);
        const float r0_2 = plom_fdivf(((P2 * (1.0f + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2);
        const float r1_2 = P10;
        const float r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        float za1, zb1;
        plom_normal_pair32(ns, &za1, &zb1);
        const float w0_2 = sqdt * za1;
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - (r0_2*X[2])*P18*w0_2;
        const float f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + + (r0_2*X[2])*P18*w0_2;
        s10 += ( + (r0_2*X[2]))*dtr + + (r0_2*X[2])*P18*w0_2;
        X[2] = plom_clipf(f2);
        const float r0_3 = plom_fdivf(((P3 * (1.0f + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3);
        const float r1_3 = P11;
        const float r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        const float w0_3 = sqdt * zb1;
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - (r0_3*X[3])*P19*w0_3;
        const float f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + + (r0_3*X[3])*P19*w0_3;
        s11 += ( + (r0_3*X[3]))*dtr + + (r0_3*X[3])*P19*w0_3;
        X[3] = plom_clipf(f3);
        X[4] = plom_clipf(f4);
        X[5] = plom_clipf(f5);
        X[6] = plom_clipf(f6);
        X[7] = plom_clipf(f7);
        const float f8 = X[8] + s8;
        X[8] = plom_clipf(f8);
        const float f9 = X[9] + s9;
        X[9] = plom_clipf(f9);
        const float f10 = X[10] + s10;
        X[10] = plom_clipf(f10);
        const float f11 = X[11] + s11;
        X[11] = plom_clipf(f11);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_no_env_sto_pp_f32(float *X, const float *pv, const float *lp, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const float P0 = lp[0];
    const float P1 = lp[1];
    const float P2 = lp[2];
    const float P3 = lp[3];
    const float P4 = lp[4];
    const float P5 = lp[5];
    const float P6 = lp[6];
    const float P7 = lp[7];
    const float P8 = lp[8];
    const float P9 = lp[9];
    const float P10 = lp[10];
    const float P11 = lp[11];
    const float P12 = lp[12];
    const float P13 = lp[13];
    const float P14 = lp[14];
    const float P15 = lp[15];
    const float P32 = PLOM_UNIFORM32(pv[32]);
    const float P33 = PLOM_UNIFORM32(pv[33]);
    const float P34 = PLOM_UNIFORM32(pv[34]);
    const float P35 = PLOM_UNIFORM32(pv[35]);
    const float P36 = PLOM_UNIFORM32(pv[36]);
    const float P37 = PLOM_UNIFORM32(pv[37]);
    const float P38 = PLOM_UNIFORM32(pv[38]);
    const float P39 = PLOM_UNIFORM32(pv[39]);
    const float P40 = PLOM_UNIFORM32(pv[40]);
    const float P41 = PLOM_UNIFORM32(pv[41]);
    const float P42 = PLOM_UNIFORM32(pv[42]);
    const float P43 = PLOM_UNIFORM32(pv[43]);
    const float P44 = PLOM_UNIFORM32(pv[44]);
    const float P45 = PLOM_UNIFORM32(pv[45]);
    const float P46 = PLOM_UNIFORM32(pv[46]);
    const float P47 = PLOM_UNIFORM32(pv[47]);
    X[8] = 0.0f; /* reset_inc */
    X[9] = 0.0f; /* reset_inc */
    X[10] = 0.0f; /* reset_inc */
    X[11] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        const float C1 = (float)plom_cov_ld(cvk + 1);
        const float C2 = (float)plom_cov_ld(cvk + 2);
        const float C3 = (float)plom_cov_ld(cvk + 3);
        const float tf0 = tft ? tft[k * 5 + 0] : plom_sinf(plom_fdivf(((2.0f * 3.141592653589793f) * tr), 365.0f));
        float s8 = 0.0f;
        float s9 = 0.0f;
        float s10 = 0.0f;
        float s11 = 0.0f;
        const float r0_0 = plom_fdivf(((P0 * (1.0f + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0);
        const float r1_0 = P8;
        const float r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        float za0, zb0;
        plom_normal_pair32(ns, &za0, &zb0);
        const float d0_0 = sqdt * za0;
        const float d1_0 = sqdt * zb0;
        float za1, zb1;
        plom_normal_pair32(ns, &za1, &zb1);
        const float d2_0 = sqdt * za1;
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr + - plom_fsqrtf((r0_0*X[0]))*d0_0+ plom_fsqrtf((r2_0))*d2_0;
        const float f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr + - plom_fsqrtf((r1_0*X[4]))*d1_0+ plom_fsqrtf((r0_0*X[0]))*d0_0;
        s8 += ( + (r0_0*X[0]))*dtr + + plom_fsqrtf((r0_0*X[0]))*d0_0;
        X[0] = plom_clipf(f0);
        const float r0_1 = plom_fdivf(((P1 * (1.0f + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1);
        const float r1_1 = P9;
        const float r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        const float d0_1 = sqdt * zb1;
        float za2, zb2;
        plom_normal_pair32(ns, &za2, &zb2);
        const float d1_1 = sqdt * za2;
        const float d2_1 = sqdt * zb2;
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr + - plom_fsqrtf((r0_1*X[1]))*d0_1+ plom_fsqrtf((r2_1))*d2_1;
        const float f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr + - plom_fsqrtf((r1_1*X[5]))*d1_1+ plom_fsqrtf((r0_1*X[1]))*d0_1;
        s9 += ( + (r0_1*X[1]))*dtr + + plom_fsqrtf((r0_1*X[1]))*d0_1;
        X[1] = plom_clipf(f1);
        const float r0_2 = plom_fdivf(((P2 * (1.0f + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2);
        const float r1_2 = P10;
        const float r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        float za3, zb3;
        plom_normal_pair32(ns, &za3, &zb3);
        const float d0_2 = sqdt * za3;
        const float d1_2 = sqdt * zb3;
        float za4, zb4;
        plom_normal_pair32(ns, &za4, &zb4);
        const float d2_2 = sqdt * za4;
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr + - plom_fsqrtf((r0_2*X[2]))*d0_2+ plom_fsqrtf((r2_2))*d2_2;
        const float f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr + - plom_fsqrtf((r1_2*X[6]))*d1_2+ plom_fsqrtf((r0_2*X[2]))*d0_2;
        s10 += ( + (r0_2*X[2]))*dtr + + plom_fsqrtf((r0_2*X[2]))*d0_2;
        X[2] = plom_clipf(f2);
        const float r0_3 = plom_fdivf(((P3 * (1.0f + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3);
        const float r1_3 = P11;
        const float r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        const float d0_3 = sqdt * zb4;
        float za5, zb5;
        plom_normal_pair32(ns, &za5, &zb5);
        const float d1_3 = sqdt * za5;
        const float d2_3 = sqdt * zb5;
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr + - plom_fsqrtf((r0_3*X[3]))*d0_3+ plom_fsqrtf((r2_3))*d2_3;
        const float f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr + - plom_fsqrtf((r1_3*X[7]))*d1_3+ plom_fsqrtf((r0_3*X[3]))*d0_3;
        s11 += ( + (r0_3*X[3]))*dtr + + plom_fsqrtf((r0_3*X[3]))*d0_3;
        X[3] = plom_clipf(f3);
        X[4] = plom_clipf(f4);
        X[5] = plom_clipf(f5);
        X[6] = plom_clipf(f6);
        X[7] = plom_clipf(f7);
        const float f8 = X[8] + s8;
        X[8] = plom_clipf(f8);
        const float f9 = X[9] + s9;
        X[9] = plom_clipf(f9);
        const float f10 = X[10] + s10;
        X[10] = plom_clipf(f10);
        const float f11 = X[11] + s11;
        X[11] = plom_clipf(f11);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_interval_no_dem_sto_no_env_sto_pp_f32(float *X, const float *pv, const float *lp, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
{
    const float dtr = (float)dt;
    (void)sqdt; (void)ns; (void)pv; (void)cv; (void)cs; (void)dm; (void)tft; (void)lp;
    const float P0 = lp[0];
    const float P1 = lp[1];
    const float P2 = lp[2];
    const float P3 = lp[3];
    const float P4 = lp[4];
    const float P5 = lp[5];
    const float P6 = lp[6];
    const float P7 = lp[7];
    const float P8 = lp[8];
    const float P9 = lp[9];
    const float P10 = lp[10];
    const float P11 = lp[11];
    const float P12 = lp[12];
    const float P13 = lp[13];
    const float P14 = lp[14];
    const float P15 = lp[15];
    const float P32 = PLOM_UNIFORM32(pv[32]);
    const float P33 = PLOM_UNIFORM32(pv[33]);
    const float P34 = PLOM_UNIFORM32(pv[34]);
    const float P35 = PLOM_UNIFORM32(pv[35]);
    const float P36 = PLOM_UNIFORM32(pv[36]);
    const float P37 = PLOM_UNIFORM32(pv[37]);
    const float P38 = PLOM_UNIFORM32(pv[38]);
    const float P39 = PLOM_UNIFORM32(pv[39]);
    const float P40 = PLOM_UNIFORM32(pv[40]);
    const float P41 = PLOM_UNIFORM32(pv[41]);
    const float P42 = PLOM_UNIFORM32(pv[42]);
    const float P43 = PLOM_UNIFORM32(pv[43]);
    const float P44 = PLOM_UNIFORM32(pv[44]);
    const float P45 = PLOM_UNIFORM32(pv[45]);
    const float P46 = PLOM_UNIFORM32(pv[46]);
    const float P47 = PLOM_UNIFORM32(pv[47]);
    X[8] = 0.0f; /* reset_inc */
    X[9] = 0.0f; /* reset_inc */
    X[10] = 0.0f; /* reset_inc */
    X[11] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        const float C1 = (float)plom_cov_ld(cvk + 1);
        const float C2 = (float)plom_cov_ld(cvk + 2);
        const float C3 = (float)plom_cov_ld(cvk + 3);
        const float tf0 = tft ? tft[k * 5 + 0] : plom_sinf(plom_fdivf(((2.0f * 3.141592653589793f) * tr), 365.0f));
        float s8 = 0.0f;
        float s9 = 0.0f;
        float s10 = 0.0f;
        float s11 = 0.0f;
        const float r0_0 = plom_fdivf(((P0 * (1.0f + (P4 * tf0))) * (P32 * X[4] + P33 * X[5] + P34 * X[6] + P35 * X[7])), C0);
        const float r1_0 = P8;
        const float r2_0 = (P12 * ((C0 - X[0]) - X[4]));
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r2_0))*dtr;
        const float f4 = X[4] + ( - (r1_0*X[4]) + (r0_0*X[0]))*dtr;
        s8 += ( + (r0_0*X[0]))*dtr;
        X[0] = plom_clipf(f0);
        const float r0_1 = plom_fdivf(((P1 * (1.0f + (P5 * tf0))) * (P36 * X[4] + P37 * X[5] + P38 * X[6] + P39 * X[7])), C1);
        const float r1_1 = P9;
        const float r2_1 = (P13 * ((C1 - X[1]) - X[5]));
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r2_1))*dtr;
        const float f5 = X[5] + ( - (r1_1*X[5]) + (r0_1*X[1]))*dtr;
        s9 += ( + (r0_1*X[1]))*dtr;
        X[1] = plom_clipf(f1);
        const float r0_2 = plom_fdivf(((P2 * (1.0f + (P6 * tf0))) * (P40 * X[4] + P41 * X[5] + P42 * X[6] + P43 * X[7])), C2);
        const float r1_2 = P10;
        const float r2_2 = (P14 * ((C2 - X[2]) - X[6]));
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r2_2))*dtr;
        const float f6 = X[6] + ( - (r1_2*X[6]) + (r0_2*X[2]))*dtr;
        s10 += ( + (r0_2*X[2]))*dtr;
        X[2] = plom_clipf(f2);
        const float r0_3 = plom_fdivf(((P3 * (1.0f + (P7 * tf0))) * (P44 * X[4] + P45 * X[5] + P46 * X[6] + P47 * X[7])), C3);
        const float r1_3 = P11;
        const float r2_3 = (P15 * ((C3 - X[3]) - X[7]));
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r2_3))*dtr;
        const float f7 = X[7] + ( - (r1_3*X[7]) + (r0_3*X[3]))*dtr;
        s11 += ( + (r0_3*X[3]))*dtr;
        X[3] = plom_clipf(f3);
        X[4] = plom_clipf(f4);
        X[5] = plom_clipf(f5);
        X[6] = plom_clipf(f6);
        X[7] = plom_clipf(f7);
        const float f8 = X[8] + s8;
        X[8] = plom_clipf(f8);
        const float f9 = X[9] + s9;
        X[9] = plom_clipf(f9);
        const float f10 = X[10] + s10;
        X[10] = plom_clipf(f10);
        const float f11 = X[11] + s11;
        X[11] = plom_clipf(f11);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_sirs4_proj2obs_pp_f32(const float *X, const double *pvd, const double *lpd, const double *cvd, double *obs)
{
    (void)pvd; (void)lpd; (void)cvd; (void)X;
    obs[0] = (double)X[8];
    obs[1] = (double)X[9];
    obs[2] = (double)X[10];
    obs[3] = (double)X[11];
}

static __device__ inline __attribute__((always_inline)) double plom_sirs4_loglik_pp(const double *obs, const double *y, const double *pvd, const double *lpd, const double *cvd, double t, double like_min, int *n_floor)
{
    double loglike = 0.0;
    int nf = 0;
    (void)t; (void)pvd; (void)lpd; (void)cvd;
    if (!plom_isnan(y[0])) {
        const double x = obs[0];
        const double gsl_mu = (lpd[20] * x);
        const double gsl_sd = plom_sqrt((((lpd[20] * (1.0 - lpd[20])) * x) + plom_powi(((lpd[20] * lpd[24]) * x), 2)));
        double like;
        if (y[0] > 0.0) like = plom_cdf_gaussian_P(y[0] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[0] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[0] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[1])) {
        const double x = obs[1];
        const double gsl_mu = (lpd[21] * x);
        const double gsl_sd = plom_sqrt((((lpd[21] * (1.0 - lpd[21])) * x) + plom_powi(((lpd[21] * lpd[25]) * x), 2)));
        double like;
        if (y[1] > 0.0) like = plom_cdf_gaussian_P(y[1] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[1] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[1] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[2])) {
        const double x = obs[2];
        const double gsl_mu = (lpd[22] * x);
        const double gsl_sd = plom_sqrt((((lpd[22] * (1.0 - lpd[22])) * x) + plom_powi(((lpd[22] * lpd[26]) * x), 2)));
        double like;
        if (y[2] > 0.0) like = plom_cdf_gaussian_P(y[2] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[2] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[2] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[3])) {
        const double x = obs[3];
        const double gsl_mu = (lpd[23] * x);
        const double gsl_sd = plom_sqrt((((lpd[23] * (1.0 - lpd[23])) * x) + plom_powi(((lpd[23] * lpd[27]) * x), 2)));
        double like;
        if (y[3] > 0.0) like = plom_cdf_gaussian_P(y[3] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[3] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[3] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    *n_floor = nf;
    return loglike;
}

struct plom_model_sirs4 {
    static constexpr int S = 12, N_TS = 4, INC_OFFSET = 8, N_INC = 4, PV = 48, N_CAC = 4, N_HAT = 16, NPP = 28, FIX_OFF = 28, N_FIX = 4, N_DRIFT = 0;
    static constexpr int NTF = 5;  /* time-only functions and covariate reciprocals of a substep */
    template <int V> static __device__ inline __attribute__((always_inline)) void interval(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
    {
        if constexpr (V == 0) plom_sirs4_interval_full_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 1) plom_sirs4_interval_no_dem_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 2) plom_sirs4_interval_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 3) plom_sirs4_interval_no_dem_sto_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 4) plom_sirs4_interval_psr_full_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 5) plom_sirs4_interval_psr_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
    }
    template <int V> static __device__ inline __attribute__((always_inline)) void interval_pp(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
    {
        if constexpr (V == 0) plom_sirs4_interval_full_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 1) plom_sirs4_interval_no_dem_sto_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 2) plom_sirs4_interval_no_env_sto_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 3) plom_sirs4_interval_no_dem_sto_no_env_sto_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 4) plom_sirs4_interval_psr_full_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 5) plom_sirs4_interval_psr_no_env_sto_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
    }
    static __device__ inline __attribute__((always_inline)) void time_funcs(double tr, const double *cvk, double *out) { plom_sirs4_time_funcs_f64(tr, cvk, out); }
    static __device__ inline __attribute__((always_inline)) void proj2obs(const double *X, const double *pvd, const double *cvd, double *obs) { plom_sirs4_proj2obs_f64(X, pvd, cvd, obs); }
    static __device__ inline __attribute__((always_inline)) void proj2obs_pp(const double *X, const double *pvd, const double *lpd, const double *cvd, double *obs) { plom_sirs4_proj2obs_pp_f64(X, pvd, lpd, cvd, obs); }
    template <int V> static __device__ inline __attribute__((always_inline)) void interval(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
    {
        if constexpr (V == 0) plom_sirs4_interval_full_f32(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 1) plom_sirs4_interval_no_dem_sto_f32(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 2) plom_sirs4_interval_no_env_sto_f32(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 3) plom_sirs4_interval_no_dem_sto_no_env_sto_f32(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
    }
    template <int V> static __device__ inline __attribute__((always_inline)) void interval_pp(float *X, const float *pv, const float *lp, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
    {
        if constexpr (V == 0) plom_sirs4_interval_full_pp_f32(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 1) plom_sirs4_interval_no_dem_sto_pp_f32(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 2) plom_sirs4_interval_no_env_sto_pp_f32(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 3) plom_sirs4_interval_no_dem_sto_no_env_sto_pp_f32(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
    }
    static __device__ inline __attribute__((always_inline)) void time_funcs(float tr, const double *cvk, float *out) { plom_sirs4_time_funcs_f32(tr, cvk, out); }
    static __device__ inline __attribute__((always_inline)) void proj2obs(const float *X, const double *pvd, const double *cvd, double *obs) { plom_sirs4_proj2obs_f32(X, pvd, cvd, obs); }
    static __device__ inline __attribute__((always_inline)) void proj2obs_pp(const float *X, const double *pvd, const double *lpd, const double *cvd, double *obs) { plom_sirs4_proj2obs_pp_f32(X, pvd, lpd, cvd, obs); }
    static __device__ inline __attribute__((always_inline)) int interval_ode(double *X, const double *pv, const double *cv, const plom_spline_set *sp, double t0, double t1, double dt, double ea, double er) { return plom_sirs4_interval_ode_f64(X, pv, cv, sp, t0, t1, dt, ea, er); }
    static __device__ inline __attribute__((always_inline)) int interval_ode_pp(double *X, const double *pv, const double *lp, const double *cv, const plom_spline_set *sp, double t0, double t1, double dt, double ea, double er) { return plom_sirs4_interval_ode_pp_f64(X, pv, lp, cv, sp, t0, t1, dt, ea, er); }
    static __device__ inline __attribute__((always_inline)) double loglik(const double *obs, const double *y, const double *pvd, const double *cvd, double t, double like_min, int *n_floor) { return plom_sirs4_loglik(obs, y, pvd, cvd, t, like_min, n_floor); }
    static __device__ inline __attribute__((always_inline)) double loglik_pp(const double *obs, const double *y, const double *pvd, const double *lpd, const double *cvd, double t, double like_min, int *n_floor) { return plom_sirs4_loglik_pp(obs, y, pvd, lpd, cvd, t, like_min, n_floor); }
    static __device__ inline __attribute__((always_inline)) void hat_values(const double *X, const double *pvd, const double *cvd, const plom_drift_view *dm, double t, double *out) { plom_sirs4_hat_values(X, pvd, cvd, dm, t, out); }
    static __device__ inline __attribute__((always_inline)) void obs_moments(const double *obs, const double *pvd, const double *cvd, double t, double *mean, double *var) { plom_sirs4_obs_moments(obs, pvd, cvd, t, mean, var); }
};
#endif

#endif
