/* Generated by iterated-inference-pipeline_amd/codegen from the seir10 model files. Do not edit. */
#ifndef PLOM_SEIR10_H
#define PLOM_SEIR10_H
#include "plom_model_common.h"

#define PLOM_SEIR10_N_CAC 10
#define PLOM_SEIR10_N_PAR_SV 3
#define PLOM_SEIR10_N_TS 10
#define PLOM_SEIR10_N_TS_INC_UNIQUE 10
#define PLOM_SEIR10_INC_OFFSET 30
#define PLOM_SEIR10_SIZE_PROJ 40
#define PLOM_SEIR10_PV_SIZE 180
#define PLOM_SEIR10_N_DEM 4
#define PLOM_SEIR10_N_ENV 1
#define PLOM_SEIR10_N_HAT 50
#define PLOM_SEIR10_N_PP 70
#define PLOM_SEIR10_FIX_OFF 70
#define PLOM_SEIR10_N_FIX 10
#define PLOM_SEIR10_N_DRIFT 0
#define PLOM_SEIR10_DRIFT_OFFSET 30

PLOM_HD void plom_seir10_time_funcs_f64(double tr, double *out)
{
    (void)tr; (void)out;
}

PLOM_HD void plom_seir10_interval_full_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
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
    const double P20 = PLOM_UNIFORM64(pv[20]);
    const double P21 = PLOM_UNIFORM64(pv[21]);
    const double P22 = PLOM_UNIFORM64(pv[22]);
    const double P23 = PLOM_UNIFORM64(pv[23]);
    const double P24 = PLOM_UNIFORM64(pv[24]);
    const double P25 = PLOM_UNIFORM64(pv[25]);
    const double P26 = PLOM_UNIFORM64(pv[26]);
    const double P27 = PLOM_UNIFORM64(pv[27]);
    const double P28 = PLOM_UNIFORM64(pv[28]);
    const double P29 = PLOM_UNIFORM64(pv[29]);
    const double P30 = PLOM_UNIFORM64(pv[30]);
    const double P31 = PLOM_UNIFORM64(pv[31]);
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
    const double P48 = PLOM_UNIFORM64(pv[48]);
    const double P49 = PLOM_UNIFORM64(pv[49]);
    const double P80 = PLOM_UNIFORM64(pv[80]);
    const double P81 = PLOM_UNIFORM64(pv[81]);
    const double P82 = PLOM_UNIFORM64(pv[82]);
    const double P83 = PLOM_UNIFORM64(pv[83]);
    const double P84 = PLOM_UNIFORM64(pv[84]);
    const double P85 = PLOM_UNIFORM64(pv[85]);
    const double P86 = PLOM_UNIFORM64(pv[86]);
    const double P87 = PLOM_UNIFORM64(pv[87]);
    const double P88 = PLOM_UNIFORM64(pv[88]);
    const double P89 = PLOM_UNIFORM64(pv[89]);
    const double P90 = PLOM_UNIFORM64(pv[90]);
    const double P91 = PLOM_UNIFORM64(pv[91]);
    const double P92 = PLOM_UNIFORM64(pv[92]);
    const double P93 = PLOM_UNIFORM64(pv[93]);
    const double P94 = PLOM_UNIFORM64(pv[94]);
    const double P95 = PLOM_UNIFORM64(pv[95]);
    const double P96 = PLOM_UNIFORM64(pv[96]);
    const double P97 = PLOM_UNIFORM64(pv[97]);
    const double P98 = PLOM_UNIFORM64(pv[98]);
    const double P99 = PLOM_UNIFORM64(pv[99]);
    const double P100 = PLOM_UNIFORM64(pv[100]);
    const double P101 = PLOM_UNIFORM64(pv[101]);
    const double P102 = PLOM_UNIFORM64(pv[102]);
    const double P103 = PLOM_UNIFORM64(pv[103]);
    const double P104 = PLOM_UNIFORM64(pv[104]);
    const double P105 = PLOM_UNIFORM64(pv[105]);
    const double P106 = PLOM_UNIFORM64(pv[106]);
    const double P107 = PLOM_UNIFORM64(pv[107]);
    const double P108 = PLOM_UNIFORM64(pv[108]);
    const double P109 = PLOM_UNIFORM64(pv[109]);
    const double P110 = PLOM_UNIFORM64(pv[110]);
    const double P111 = PLOM_UNIFORM64(pv[111]);
    const double P112 = PLOM_UNIFORM64(pv[112]);
    const double P113 = PLOM_UNIFORM64(pv[113]);
    const double P114 = PLOM_UNIFORM64(pv[114]);
    const double P115 = PLOM_UNIFORM64(pv[115]);
    const double P116 = PLOM_UNIFORM64(pv[116]);
    const double P117 = PLOM_UNIFORM64(pv[117]);
    const double P118 = PLOM_UNIFORM64(pv[118]);
    const double P119 = PLOM_UNIFORM64(pv[119]);
    const double P120 = PLOM_UNIFORM64(pv[120]);
    const double P121 = PLOM_UNIFORM64(pv[121]);
    const double P122 = PLOM_UNIFORM64(pv[122]);
    const double P123 = PLOM_UNIFORM64(pv[123]);
    const double P124 = PLOM_UNIFORM64(pv[124]);
    const double P125 = PLOM_UNIFORM64(pv[125]);
    const double P126 = PLOM_UNIFORM64(pv[126]);
    const double P127 = PLOM_UNIFORM64(pv[127]);
    const double P128 = PLOM_UNIFORM64(pv[128]);
    const double P129 = PLOM_UNIFORM64(pv[129]);
    const double P130 = PLOM_UNIFORM64(pv[130]);
    const double P131 = PLOM_UNIFORM64(pv[131]);
    const double P132 = PLOM_UNIFORM64(pv[132]);
    const double P133 = PLOM_UNIFORM64(pv[133]);
    const double P134 = PLOM_UNIFORM64(pv[134]);
    const double P135 = PLOM_UNIFORM64(pv[135]);
    const double P136 = PLOM_UNIFORM64(pv[136]);
    const double P137 = PLOM_UNIFORM64(pv[137]);
    const double P138 = PLOM_UNIFORM64(pv[138]);
    const double P139 = PLOM_UNIFORM64(pv[139]);
    const double P140 = PLOM_UNIFORM64(pv[140]);
    const double P141 = PLOM_UNIFORM64(pv[141]);
    const double P142 = PLOM_UNIFORM64(pv[142]);
    const double P143 = PLOM_UNIFORM64(pv[143]);
    const double P144 = PLOM_UNIFORM64(pv[144]);
    const double P145 = PLOM_UNIFORM64(pv[145]);
    const double P146 = PLOM_UNIFORM64(pv[146]);
    const double P147 = PLOM_UNIFORM64(pv[147]);
    const double P148 = PLOM_UNIFORM64(pv[148]);
    const double P149 = PLOM_UNIFORM64(pv[149]);
    const double P150 = PLOM_UNIFORM64(pv[150]);
    const double P151 = PLOM_UNIFORM64(pv[151]);
    const double P152 = PLOM_UNIFORM64(pv[152]);
    const double P153 = PLOM_UNIFORM64(pv[153]);
    const double P154 = PLOM_UNIFORM64(pv[154]);
    const double P155 = PLOM_UNIFORM64(pv[155]);
    const double P156 = PLOM_UNIFORM64(pv[156]);
    const double P157 = PLOM_UNIFORM64(pv[157]);
    const double P158 = PLOM_UNIFORM64(pv[158]);
    const double P159 = PLOM_UNIFORM64(pv[159]);
    const double P160 = PLOM_UNIFORM64(pv[160]);
    const double P161 = PLOM_UNIFORM64(pv[161]);
    const double P162 = PLOM_UNIFORM64(pv[162]);
    const double P163 = PLOM_UNIFORM64(pv[163]);
    const double P164 = PLOM_UNIFORM64(pv[164]);
    const double P165 = PLOM_UNIFORM64(pv[165]);
    const double P166 = PLOM_UNIFORM64(pv[166]);
    const double P167 = PLOM_UNIFORM64(pv[167]);
    const double P168 = PLOM_UNIFORM64(pv[168]);
    const double P169 = PLOM_UNIFORM64(pv[169]);
    const double P170 = PLOM_UNIFORM64(pv[170]);
    const double P171 = PLOM_UNIFORM64(pv[171]);
    const double P172 = PLOM_UNIFORM64(pv[172]);
    const double P173 = PLOM_UNIFORM64(pv[173]);
    const double P174 = PLOM_UNIFORM64(pv[174]);
    const double P175 = PLOM_UNIFORM64(pv[175]);
    const double P176 = PLOM_UNIFORM64(pv[176]);
    const double P177 = PLOM_UNIFORM64(pv[177]);
    const double P178 = PLOM_UNIFORM64(pv[178]);
    const double P179 = PLOM_UNIFORM64(pv[179]);
    X[30] = 0.0; /* reset_inc */
    X[31] = 0.0; /* reset_inc */
    X[32] = 0.0; /* reset_inc */
    X[33] = 0.0; /* reset_inc */
    X[34] = 0.0; /* reset_inc */
    X[35] = 0.0; /* reset_inc */
    X[36] = 0.0; /* reset_inc */
    X[37] = 0.0; /* reset_inc */
    X[38] = 0.0; /* reset_inc */
    X[39] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double C4 = plom_cov_ld(cvk + 4);
        const double C5 = plom_cov_ld(cvk + 5);
        const double C6 = plom_cov_ld(cvk + 6);
        const double C7 = plom_cov_ld(cvk + 7);
        const double C8 = plom_cov_ld(cvk + 8);
        const double C9 = plom_cov_ld(cvk + 9);
        double s30 = 0.0;
        double s31 = 0.0;
        double s32 = 0.0;
        double s33 = 0.0;
        double s34 = 0.0;
        double s35 = 0.0;
        double s36 = 0.0;
        double s37 = 0.0;
        double s38 = 0.0;
        double s39 = 0.0;
        const double r0_0 = ((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])) / C0);
        const double r1_0 = P10;
        const double r2_0 = P20;
        const double r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        double za0, zb0;
        plom_normal_pair64(ns, &za0, &zb0);
        const double d0_0 = sqdt * za0;
        const double d1_0 = sqdt * zb0;
        double za1, zb1;
        plom_normal_pair64(ns, &za1, &zb1);
        const double d2_0 = sqdt * za1;
        const double d3_0 = sqdt * zb1;
        double za2, zb2;
        plom_normal_pair64(ns, &za2, &zb2);
        const double w0_0 = sqdt * za2;
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r3_0))*dtr + - plom_sqrt_rate((r0_0*X[0]))*d0_0+ plom_sqrt_rate((r3_0))*d3_0 + - (r0_0*X[0])*P40*w0_0;
        const double f10 = X[10] + ( - (r1_0*X[10]) + (r0_0*X[0]))*dtr + - plom_sqrt_rate((r1_0*X[10]))*d1_0+ plom_sqrt_rate((r0_0*X[0]))*d0_0 + + (r0_0*X[0])*P40*w0_0;
        const double f20 = X[20] + ( - (r2_0*X[20]) + (r1_0*X[10]))*dtr + - plom_sqrt_rate((r2_0*X[20]))*d2_0+ plom_sqrt_rate((r1_0*X[10]))*d1_0;
        s30 += ( + (r1_0*X[10]))*dtr + + plom_sqrt_rate((r1_0*X[10]))*d1_0;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        const double r0_1 = ((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])) / C1);
        const double r1_1 = P11;
        const double r2_1 = P21;
        const double r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        const double d0_1 = sqdt * zb2;
        double za3, zb3;
        plom_normal_pair64(ns, &za3, &zb3);
        const double d1_1 = sqdt * za3;
        const double d2_1 = sqdt * zb3;
        double za4, zb4;
        plom_normal_pair64(ns, &za4, &zb4);
        const double d3_1 = sqdt * za4;
        const double w0_1 = sqdt * zb4;
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r3_1))*dtr + - plom_sqrt_rate((r0_1*X[1]))*d0_1+ plom_sqrt_rate((r3_1))*d3_1 + - (r0_1*X[1])*P41*w0_1;
        const double f11 = X[11] + ( - (r1_1*X[11]) + (r0_1*X[1]))*dtr + - plom_sqrt_rate((r1_1*X[11]))*d1_1+ plom_sqrt_rate((r0_1*X[1]))*d0_1 + + (r0_1*X[1])*P41*w0_1;
        const double f21 = X[21] + ( - (r2_1*X[21]) + (r1_1*X[11]))*dtr + - plom_sqrt_rate((r2_1*X[21]))*d2_1+ plom_sqrt_rate((r1_1*X[11]))*d1_1;
        s31 += ( + (r1_1*X[11]))*dtr + + plom_sqrt_rate((r1_1*X[11]))*d1_1;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        const double r0_2 = ((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])) / C2);
        const double r1_2 = P12;
        const double r2_2 = P22;
        const double r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        double za5, zb5;
        plom_normal_pair64(ns, &za5, &zb5);
        const double d0_2 = sqdt * za5;
        const double d1_2 = sqdt * zb5;
        double za6, zb6;
        plom_normal_pair64(ns, &za6, &zb6);
        const double d2_2 = sqdt * za6;
        const double d3_2 = sqdt * zb6;
        double za7, zb7;
        plom_normal_pair64(ns, &za7, &zb7);
        const double w0_2 = sqdt * za7;
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r3_2))*dtr + - plom_sqrt_rate((r0_2*X[2]))*d0_2+ plom_sqrt_rate((r3_2))*d3_2 + - (r0_2*X[2])*P42*w0_2;
        const double f12 = X[12] + ( - (r1_2*X[12]) + (r0_2*X[2]))*dtr + - plom_sqrt_rate((r1_2*X[12]))*d1_2+ plom_sqrt_rate((r0_2*X[2]))*d0_2 + + (r0_2*X[2])*P42*w0_2;
        const double f22 = X[22] + ( - (r2_2*X[22]) + (r1_2*X[12]))*dtr + - plom_sqrt_rate((r2_2*X[22]))*d2_2+ plom_sqrt_rate((r1_2*X[12]))*d1_2;
        s32 += ( + (r1_2*X[12]))*dtr + + plom_sqrt_rate((r1_2*X[12]))*d1_2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        X[12] = (f12 < 0.0) ? 0.0 : f12;
        const double r0_3 = ((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])) / C3);
        const double r1_3 = P13;
        const double r2_3 = P23;
        const double r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        const double d0_3 = sqdt * zb7;
        double za8, zb8;
        plom_normal_pair64(ns, &za8, &zb8);
        const double d1_3 = sqdt * za8;
        const double d2_3 = sqdt * zb8;
        double za9, zb9;
        plom_normal_pair64(ns, &za9, &zb9);
        const double d3_3 = sqdt * za9;
        const double w0_3 = sqdt * zb9;
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r3_3))*dtr + - plom_sqrt_rate((r0_3*X[3]))*d0_3+ plom_sqrt_rate((r3_3))*d3_3 + - (r0_3*X[3])*P43*w0_3;
        const double f13 = X[13] + ( - (r1_3*X[13]) + (r0_3*X[3]))*dtr + - plom_sqrt_rate((r1_3*X[13]))*d1_3+ plom_sqrt_rate((r0_3*X[3]))*d0_3 + + (r0_3*X[3])*P43*w0_3;
        const double f23 = X[23] + ( - (r2_3*X[23]) + (r1_3*X[13]))*dtr + - plom_sqrt_rate((r2_3*X[23]))*d2_3+ plom_sqrt_rate((r1_3*X[13]))*d1_3;
        s33 += ( + (r1_3*X[13]))*dtr + + plom_sqrt_rate((r1_3*X[13]))*d1_3;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[13] = (f13 < 0.0) ? 0.0 : f13;
        const double r0_4 = ((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])) / C4);
        const double r1_4 = P14;
        const double r2_4 = P24;
        const double r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        double za10, zb10;
        plom_normal_pair64(ns, &za10, &zb10);
        const double d0_4 = sqdt * za10;
        const double d1_4 = sqdt * zb10;
        double za11, zb11;
        plom_normal_pair64(ns, &za11, &zb11);
        const double d2_4 = sqdt * za11;
        const double d3_4 = sqdt * zb11;
        double za12, zb12;
        plom_normal_pair64(ns, &za12, &zb12);
        const double w0_4 = sqdt * za12;
        const double f4 = X[4] + ( - (r0_4*X[4]) + (r3_4))*dtr + - plom_sqrt_rate((r0_4*X[4]))*d0_4+ plom_sqrt_rate((r3_4))*d3_4 + - (r0_4*X[4])*P44*w0_4;
        const double f14 = X[14] + ( - (r1_4*X[14]) + (r0_4*X[4]))*dtr + - plom_sqrt_rate((r1_4*X[14]))*d1_4+ plom_sqrt_rate((r0_4*X[4]))*d0_4 + + (r0_4*X[4])*P44*w0_4;
        const double f24 = X[24] + ( - (r2_4*X[24]) + (r1_4*X[14]))*dtr + - plom_sqrt_rate((r2_4*X[24]))*d2_4+ plom_sqrt_rate((r1_4*X[14]))*d1_4;
        s34 += ( + (r1_4*X[14]))*dtr + + plom_sqrt_rate((r1_4*X[14]))*d1_4;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[14] = (f14 < 0.0) ? 0.0 : f14;
        const double r0_5 = ((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])) / C5);
        const double r1_5 = P15;
        const double r2_5 = P25;
        const double r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        const double d0_5 = sqdt * zb12;
        double za13, zb13;
        plom_normal_pair64(ns, &za13, &zb13);
        const double d1_5 = sqdt * za13;
        const double d2_5 = sqdt * zb13;
        double za14, zb14;
        plom_normal_pair64(ns, &za14, &zb14);
        const double d3_5 = sqdt * za14;
        const double w0_5 = sqdt * zb14;
        const double f5 = X[5] + ( - (r0_5*X[5]) + (r3_5))*dtr + - plom_sqrt_rate((r0_5*X[5]))*d0_5+ plom_sqrt_rate((r3_5))*d3_5 + - (r0_5*X[5])*P45*w0_5;
        const double f15 = X[15] + ( - (r1_5*X[15]) + (r0_5*X[5]))*dtr + - plom_sqrt_rate((r1_5*X[15]))*d1_5+ plom_sqrt_rate((r0_5*X[5]))*d0_5 + + (r0_5*X[5])*P45*w0_5;
        const double f25 = X[25] + ( - (r2_5*X[25]) + (r1_5*X[15]))*dtr + - plom_sqrt_rate((r2_5*X[25]))*d2_5+ plom_sqrt_rate((r1_5*X[15]))*d1_5;
        s35 += ( + (r1_5*X[15]))*dtr + + plom_sqrt_rate((r1_5*X[15]))*d1_5;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[15] = (f15 < 0.0) ? 0.0 : f15;
        const double r0_6 = ((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])) / C6);
        const double r1_6 = P16;
        const double r2_6 = P26;
        const double r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        double za15, zb15;
        plom_normal_pair64(ns, &za15, &zb15);
        const double d0_6 = sqdt * za15;
        const double d1_6 = sqdt * zb15;
        double za16, zb16;
        plom_normal_pair64(ns, &za16, &zb16);
        const double d2_6 = sqdt * za16;
        const double d3_6 = sqdt * zb16;
        double za17, zb17;
        plom_normal_pair64(ns, &za17, &zb17);
        const double w0_6 = sqdt * za17;
        const double f6 = X[6] + ( - (r0_6*X[6]) + (r3_6))*dtr + - plom_sqrt_rate((r0_6*X[6]))*d0_6+ plom_sqrt_rate((r3_6))*d3_6 + - (r0_6*X[6])*P46*w0_6;
        const double f16 = X[16] + ( - (r1_6*X[16]) + (r0_6*X[6]))*dtr + - plom_sqrt_rate((r1_6*X[16]))*d1_6+ plom_sqrt_rate((r0_6*X[6]))*d0_6 + + (r0_6*X[6])*P46*w0_6;
        const double f26 = X[26] + ( - (r2_6*X[26]) + (r1_6*X[16]))*dtr + - plom_sqrt_rate((r2_6*X[26]))*d2_6+ plom_sqrt_rate((r1_6*X[16]))*d1_6;
        s36 += ( + (r1_6*X[16]))*dtr + + plom_sqrt_rate((r1_6*X[16]))*d1_6;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[16] = (f16 < 0.0) ? 0.0 : f16;
        const double r0_7 = ((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])) / C7);
        const double r1_7 = P17;
        const double r2_7 = P27;
        const double r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        const double d0_7 = sqdt * zb17;
        double za18, zb18;
        plom_normal_pair64(ns, &za18, &zb18);
        const double d1_7 = sqdt * za18;
        const double d2_7 = sqdt * zb18;
        double za19, zb19;
        plom_normal_pair64(ns, &za19, &zb19);
        const double d3_7 = sqdt * za19;
        const double w0_7 = sqdt * zb19;
        const double f7 = X[7] + ( - (r0_7*X[7]) + (r3_7))*dtr + - plom_sqrt_rate((r0_7*X[7]))*d0_7+ plom_sqrt_rate((r3_7))*d3_7 + - (r0_7*X[7])*P47*w0_7;
        const double f17 = X[17] + ( - (r1_7*X[17]) + (r0_7*X[7]))*dtr + - plom_sqrt_rate((r1_7*X[17]))*d1_7+ plom_sqrt_rate((r0_7*X[7]))*d0_7 + + (r0_7*X[7])*P47*w0_7;
        const double f27 = X[27] + ( - (r2_7*X[27]) + (r1_7*X[17]))*dtr + - plom_sqrt_rate((r2_7*X[27]))*d2_7+ plom_sqrt_rate((r1_7*X[17]))*d1_7;
        s37 += ( + (r1_7*X[17]))*dtr + + plom_sqrt_rate((r1_7*X[17]))*d1_7;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        X[17] = (f17 < 0.0) ? 0.0 : f17;
        const double r0_8 = ((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])) / C8);
        const double r1_8 = P18;
        const double r2_8 = P28;
        const double r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        double za20, zb20;
        plom_normal_pair64(ns, &za20, &zb20);
        const double d0_8 = sqdt * za20;
        const double d1_8 = sqdt * zb20;
        double za21, zb21;
        plom_normal_pair64(ns, &za21, &zb21);
        const double d2_8 = sqdt * za21;
        const double d3_8 = sqdt * zb21;
        double za22, zb22;
        plom_normal_pair64(ns, &za22, &zb22);
        const double w0_8 = sqdt * za22;
        const double f8 = X[8] + ( - (r0_8*X[8]) + (r3_8))*dtr + - plom_sqrt_rate((r0_8*X[8]))*d0_8+ plom_sqrt_rate((r3_8))*d3_8 + - (r0_8*X[8])*P48*w0_8;
        const double f18 = X[18] + ( - (r1_8*X[18]) + (r0_8*X[8]))*dtr + - plom_sqrt_rate((r1_8*X[18]))*d1_8+ plom_sqrt_rate((r0_8*X[8]))*d0_8 + + (r0_8*X[8])*P48*w0_8;
        const double f28 = X[28] + ( - (r2_8*X[28]) + (r1_8*X[18]))*dtr + - plom_sqrt_rate((r2_8*X[28]))*d2_8+ plom_sqrt_rate((r1_8*X[18]))*d1_8;
        s38 += ( + (r1_8*X[18]))*dtr + + plom_sqrt_rate((r1_8*X[18]))*d1_8;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        X[18] = (f18 < 0.0) ? 0.0 : f18;
        const double r0_9 = ((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])) / C9);
        const double r1_9 = P19;
        const double r2_9 = P29;
        const double r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        const double d0_9 = sqdt * zb22;
        double za23, zb23;
        plom_normal_pair64(ns, &za23, &zb23);
        const double d1_9 = sqdt * za23;
        const double d2_9 = sqdt * zb23;
        double za24, zb24;
        plom_normal_pair64(ns, &za24, &zb24);
        const double d3_9 = sqdt * za24;
        const double w0_9 = sqdt * zb24;
        const double f9 = X[9] + ( - (r0_9*X[9]) + (r3_9))*dtr + - plom_sqrt_rate((r0_9*X[9]))*d0_9+ plom_sqrt_rate((r3_9))*d3_9 + - (r0_9*X[9])*P49*w0_9;
        const double f19 = X[19] + ( - (r1_9*X[19]) + (r0_9*X[9]))*dtr + - plom_sqrt_rate((r1_9*X[19]))*d1_9+ plom_sqrt_rate((r0_9*X[9]))*d0_9 + + (r0_9*X[9])*P49*w0_9;
        const double f29 = X[29] + ( - (r2_9*X[29]) + (r1_9*X[19]))*dtr + - plom_sqrt_rate((r2_9*X[29]))*d2_9+ plom_sqrt_rate((r1_9*X[19]))*d1_9;
        s39 += ( + (r1_9*X[19]))*dtr + + plom_sqrt_rate((r1_9*X[19]))*d1_9;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        X[19] = (f19 < 0.0) ? 0.0 : f19;
        X[20] = (f20 < 0.0) ? 0.0 : f20;
        X[21] = (f21 < 0.0) ? 0.0 : f21;
        X[22] = (f22 < 0.0) ? 0.0 : f22;
        X[23] = (f23 < 0.0) ? 0.0 : f23;
        X[24] = (f24 < 0.0) ? 0.0 : f24;
        X[25] = (f25 < 0.0) ? 0.0 : f25;
        X[26] = (f26 < 0.0) ? 0.0 : f26;
        X[27] = (f27 < 0.0) ? 0.0 : f27;
        X[28] = (f28 < 0.0) ? 0.0 : f28;
        X[29] = (f29 < 0.0) ? 0.0 : f29;
        const double f30 = X[30] + s30;
        X[30] = (f30 < 0.0) ? 0.0 : f30;
        const double f31 = X[31] + s31;
        X[31] = (f31 < 0.0) ? 0.0 : f31;
        const double f32 = X[32] + s32;
        X[32] = (f32 < 0.0) ? 0.0 : f32;
        const double f33 = X[33] + s33;
        X[33] = (f33 < 0.0) ? 0.0 : f33;
        const double f34 = X[34] + s34;
        X[34] = (f34 < 0.0) ? 0.0 : f34;
        const double f35 = X[35] + s35;
        X[35] = (f35 < 0.0) ? 0.0 : f35;
        const double f36 = X[36] + s36;
        X[36] = (f36 < 0.0) ? 0.0 : f36;
        const double f37 = X[37] + s37;
        X[37] = (f37 < 0.0) ? 0.0 : f37;
        const double f38 = X[38] + s38;
        X[38] = (f38 < 0.0) ? 0.0 : f38;
        const double f39 = X[39] + s39;
        X[39] = (f39 < 0.0) ? 0.0 : f39;
        t += dt;
    }
}

PLOM_HD void plom_seir10_interval_no_dem_sto_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
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
    const double P20 = PLOM_UNIFORM64(pv[20]);
    const double P21 = PLOM_UNIFORM64(pv[21]);
    const double P22 = PLOM_UNIFORM64(pv[22]);
    const double P23 = PLOM_UNIFORM64(pv[23]);
    const double P24 = PLOM_UNIFORM64(pv[24]);
    const double P25 = PLOM_UNIFORM64(pv[25]);
    const double P26 = PLOM_UNIFORM64(pv[26]);
    const double P27 = PLOM_UNIFORM64(pv[27]);
    const double P28 = PLOM_UNIFORM64(pv[28]);
    const double P29 = PLOM_UNIFORM64(pv[29]);
    const double P30 = PLOM_UNIFORM64(pv[30]);
    const double P31 = PLOM_UNIFORM64(pv[31]);
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
    const double P48 = PLOM_UNIFORM64(pv[48]);
    const double P49 = PLOM_UNIFORM64(pv[49]);
    const double P80 = PLOM_UNIFORM64(pv[80]);
    const double P81 = PLOM_UNIFORM64(pv[81]);
    const double P82 = PLOM_UNIFORM64(pv[82]);
    const double P83 = PLOM_UNIFORM64(pv[83]);
    const double P84 = PLOM_UNIFORM64(pv[84]);
    const double P85 = PLOM_UNIFORM64(pv[85]);
    const double P86 = PLOM_UNIFORM64(pv[86]);
    const double P87 = PLOM_UNIFORM64(pv[87]);
    const double P88 = PLOM_UNIFORM64(pv[88]);
    const double P89 = PLOM_UNIFORM64(pv[89]);
    const double P90 = PLOM_UNIFORM64(pv[90]);
    const double P91 = PLOM_UNIFORM64(pv[91]);
    const double P92 = PLOM_UNIFORM64(pv[92]);
    const double P93 = PLOM_UNIFORM64(pv[93]);
    const double P94 = PLOM_UNIFORM64(pv[94]);
    const double P95 = PLOM_UNIFORM64(pv[95]);
    const double P96 = PLOM_UNIFORM64(pv[96]);
    const double P97 = PLOM_UNIFORM64(pv[97]);
    const double P98 = PLOM_UNIFORM64(pv[98]);
    const double P99 = PLOM_UNIFORM64(pv[99]);
    const double P100 = PLOM_UNIFORM64(pv[100]);
    const double P101 = PLOM_UNIFORM64(pv[101]);
    const double P102 = PLOM_UNIFORM64(pv[102]);
    const double P103 = PLOM_UNIFORM64(pv[103]);
    const double P104 = PLOM_UNIFORM64(pv[104]);
    const double P105 = PLOM_UNIFORM64(pv[105]);
    const double P106 = PLOM_UNIFORM64(pv[106]);
    const double P107 = PLOM_UNIFORM64(pv[107]);
    const double P108 = PLOM_UNIFORM64(pv[108]);
    const double P109 = PLOM_UNIFORM64(pv[109]);
    const double P110 = PLOM_UNIFORM64(pv[110]);
    const double P111 = PLOM_UNIFORM64(pv[111]);
    const double P112 = PLOM_UNIFORM64(pv[112]);
    const double P113 = PLOM_UNIFORM64(pv[113]);
    const double P114 = PLOM_UNIFORM64(pv[114]);
    const double P115 = PLOM_UNIFORM64(pv[115]);
    const double P116 = PLOM_UNIFORM64(pv[116]);
    const double P117 = PLOM_UNIFORM64(pv[117]);
    const double P118 = PLOM_UNIFORM64(pv[118]);
    const double P119 = PLOM_UNIFORM64(pv[119]);
    const double P120 = PLOM_UNIFORM64(pv[120]);
    const double P121 = PLOM_UNIFORM64(pv[121]);
    const double P122 = PLOM_UNIFORM64(pv[122]);
    const double P123 = PLOM_UNIFORM64(pv[123]);
    const double P124 = PLOM_UNIFORM64(pv[124]);
    const double P125 = PLOM_UNIFORM64(pv[125]);
    const double P126 = PLOM_UNIFORM64(pv[126]);
    const double P127 = PLOM_UNIFORM64(pv[127]);
    const double P128 = PLOM_UNIFORM64(pv[128]);
    const double P129 = PLOM_UNIFORM64(pv[129]);
    const double P130 = PLOM_UNIFORM64(pv[130]);
    const double P131 = PLOM_UNIFORM64(pv[131]);
    const double P132 = PLOM_UNIFORM64(pv[132]);
    const double P133 = PLOM_UNIFORM64(pv[133]);
    const double P134 = PLOM_UNIFORM64(pv[134]);
    const double P135 = PLOM_UNIFORM64(pv[135]);
    const double P136 = PLOM_UNIFORM64(pv[136]);
    const double P137 = PLOM_UNIFORM64(pv[137]);
    const double P138 = PLOM_UNIFORM64(pv[138]);
    const double P139 = PLOM_UNIFORM64(pv[139]);
    const double P140 = PLOM_UNIFORM64(pv[140]);
    const double P141 = PLOM_UNIFORM64(pv[141]);
    const double P142 = PLOM_UNIFORM64(pv[142]);
    const double P143 = PLOM_UNIFORM64(pv[143]);
    const double P144 = PLOM_UNIFORM64(pv[144]);
    const double P145 = PLOM_UNIFORM64(pv[145]);
    const double P146 = PLOM_UNIFORM64(pv[146]);
    const double P147 = PLOM_UNIFORM64(pv[147]);
    const double P148 = PLOM_UNIFORM64(pv[148]);
    const double P149 = PLOM_UNIFORM64(pv[149]);
    const double P150 = PLOM_UNIFORM64(pv[150]);
    const double P151 = PLOM_UNIFORM64(pv[151]);
    const double P152 = PLOM_UNIFORM64(pv[152]);
    const double P153 = PLOM_UNIFORM64(pv[153]);
    const double P154 = PLOM_UNIFORM64(pv[154]);
    const double P155 = PLOM_UNIFORM64(pv[155]);
    const double P156 = PLOM_UNIFORM64(pv[156]);
    const double P157 = PLOM_UNIFORM64(pv[157]);
    const double P158 = PLOM_UNIFORM64(pv[158]);
    const double P159 = PLOM_UNIFORM64(pv[159]);
    const double P160 = PLOM_UNIFORM64(pv[160]);
    const double P161 = PLOM_UNIFORM64(pv[161]);
    const double P162 = PLOM_UNIFORM64(pv[162]);
    const double P163 = PLOM_UNIFORM64(pv[163]);
    const double P164 = PLOM_UNIFORM64(pv[164]);
    const double P165 = PLOM_UNIFORM64(pv[165]);
    const double P166 = PLOM_UNIFORM64(pv[166]);
    const double P167 = PLOM_UNIFORM64(pv[167]);
    const double P168 = PLOM_UNIFORM64(pv[168]);
    const double P169 = PLOM_UNIFORM64(pv[169]);
    const double P170 = PLOM_UNIFORM64(pv[170]);
    const double P171 = PLOM_UNIFORM64(pv[171]);
    const double P172 = PLOM_UNIFORM64(pv[172]);
    const double P173 = PLOM_UNIFORM64(pv[173]);
    const double P174 = PLOM_UNIFORM64(pv[174]);
    const double P175 = PLOM_UNIFORM64(pv[175]);
    const double P176 = PLOM_UNIFORM64(pv[176]);
    const double P177 = PLOM_UNIFORM64(pv[177]);
    const double P178 = PLOM_UNIFORM64(pv[178]);
    const double P179 = PLOM_UNIFORM64(pv[179]);
    X[30] = 0.0; /* reset_inc */
    X[31] = 0.0; /* reset_inc */
    X[32] = 0.0; /* reset_inc */
    X[33] = 0.0; /* reset_inc */
    X[34] = 0.0; /* reset_inc */
    X[35] = 0.0; /* reset_inc */
    X[36] = 0.0; /* reset_inc */
    X[37] = 0.0; /* reset_inc */
    X[38] = 0.0; /* reset_inc */
    X[39] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double C4 = plom_cov_ld(cvk + 4);
        const double C5 = plom_cov_ld(cvk + 5);
        const double C6 = plom_cov_ld(cvk + 6);
        const double C7 = plom_cov_ld(cvk + 7);
        const double C8 = plom_cov_ld(cvk + 8);
        const double C9 = plom_cov_ld(cvk + 9);
        double s30 = 0.0;
        double s31 = 0.0;
        double s32 = 0.0;
        double s33 = 0.0;
        double s34 = 0.0;
        double s35 = 0.0;
        double s36 = 0.0;
        double s37 = 0.0;
        double s38 = 0.0;
        double s39 = 0.0;
        const double r0_0 = ((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])) / C0);
        const double r1_0 = P10;
        const double r2_0 = P20;
        const double r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        double za0, zb0;
        plom_normal_pair64(ns, &za0, &zb0);
        const double w0_0 = sqdt * za0;
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r3_0))*dtr + - (r0_0*X[0])*P40*w0_0;
        const double f10 = X[10] + ( - (r1_0*X[10]) + (r0_0*X[0]))*dtr + + (r0_0*X[0])*P40*w0_0;
        const double f20 = X[20] + ( - (r2_0*X[20]) + (r1_0*X[10]))*dtr;
        s30 += ( + (r1_0*X[10]))*dtr;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        const double r0_1 = ((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])) / C1);
        const double r1_1 = P11;
        const double r2_1 = P21;
        const double r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        const double w0_1 = sqdt * zb0;
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r3_1))*dtr + - (r0_1*X[1])*P41*w0_1;
        const double f11 = X[11] + ( - (r1_1*X[11]) + (r0_1*X[1]))*dtr + + (r0_1*X[1])*P41*w0_1;
        const double f21 = X[21] + ( - (r2_1*X[21]) + (r1_1*X[11]))*dtr;
        s31 += ( + (r1_1*X[11]))*dtr;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        const double r0_2 = ((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])) / C2);
        const double r1_2 = P12;
        const double r2_2 = P22;
        const double r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        double za1, zb1;
        plom_normal_pair64(ns, &za1, &zb1);
        const double w0_2 = sqdt * za1;
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r3_2))*dtr + - (r0_2*X[2])*P42*w0_2;
        const double f12 = X[12] + ( - (r1_2*X[12]) + (r0_2*X[2]))*dtr + + (r0_2*X[2])*P42*w0_2;
        const double f22 = X[22] + ( - (r2_2*X[22]) + (r1_2*X[12]))*dtr;
        s32 += ( + (r1_2*X[12]))*dtr;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        X[12] = (f12 < 0.0) ? 0.0 : f12;
        const double r0_3 = ((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])) / C3);
        const double r1_3 = P13;
        const double r2_3 = P23;
        const double r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        const double w0_3 = sqdt * zb1;
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r3_3))*dtr + - (r0_3*X[3])*P43*w0_3;
        const double f13 = X[13] + ( - (r1_3*X[13]) + (r0_3*X[3]))*dtr + + (r0_3*X[3])*P43*w0_3;
        const double f23 = X[23] + ( - (r2_3*X[23]) + (r1_3*X[13]))*dtr;
        s33 += ( + (r1_3*X[13]))*dtr;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[13] = (f13 < 0.0) ? 0.0 : f13;
        const double r0_4 = ((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])) / C4);
        const double r1_4 = P14;
        const double r2_4 = P24;
        const double r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        double za2, zb2;
        plom_normal_pair64(ns, &za2, &zb2);
        const double w0_4 = sqdt * za2;
        const double f4 = X[4] + ( - (r0_4*X[4]) + (r3_4))*dtr + - (r0_4*X[4])*P44*w0_4;
        const double f14 = X[14] + ( - (r1_4*X[14]) + (r0_4*X[4]))*dtr + + (r0_4*X[4])*P44*w0_4;
        const double f24 = X[24] + ( - (r2_4*X[24]) + (r1_4*X[14]))*dtr;
        s34 += ( + (r1_4*X[14]))*dtr;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[14] = (f14 < 0.0) ? 0.0 : f14;
        const double r0_5 = ((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])) / C5);
        const double r1_5 = P15;
        const double r2_5 = P25;
        const double r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        const double w0_5 = sqdt * zb2;
        const double f5 = X[5] + ( - (r0_5*X[5]) + (r3_5))*dtr + - (r0_5*X[5])*P45*w0_5;
        const double f15 = X[15] + ( - (r1_5*X[15]) + (r0_5*X[5]))*dtr + + (r0_5*X[5])*P45*w0_5;
        const double f25 = X[25] + ( - (r2_5*X[25]) + (r1_5*X[15]))*dtr;
        s35 += ( + (r1_5*X[15]))*dtr;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[15] = (f15 < 0.0) ? 0.0 : f15;
        const double r0_6 = ((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])) / C6);
        const double r1_6 = P16;
        const double r2_6 = P26;
        const double r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        double za3, zb3;
        plom_normal_pair64(ns, &za3, &zb3);
        const double w0_6 = sqdt * za3;
        const double f6 = X[6] + ( - (r0_6*X[6]) + (r3_6))*dtr + - (r0_6*X[6])*P46*w0_6;
        const double f16 = X[16] + ( - (r1_6*X[16]) + (r0_6*X[6]))*dtr + + (r0_6*X[6])*P46*w0_6;
        const double f26 = X[26] + ( - (r2_6*X[26]) + (r1_6*X[16]))*dtr;
        s36 += ( + (r1_6*X[16]))*dtr;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[16] = (f16 < 0.0) ? 0.0 : f16;
        const double r0_7 = ((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])) / C7);
        const double r1_7 = P17;
        const double r2_7 = P27;
        const double r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        const double w0_7 = sqdt * zb3;
        const double f7 = X[7] + ( - (r0_7*X[7]) + (r3_7))*dtr + - (r0_7*X[7])*P47*w0_7;
        const double f17 = X[17] + ( - (r1_7*X[17]) + (r0_7*X[7]))*dtr + + (r0_7*X[7])*P47*w0_7;
        const double f27 = X[27] + ( - (r2_7*X[27]) + (r1_7*X[17]))*dtr;
        s37 += ( + (r1_7*X[17]))*dtr;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        X[17] = (f17 < 0.0) ? 0.0 : f17;
        const double r0_8 = ((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])) / C8);
        const double r1_8 = P18;
        const double r2_8 = P28;
        const double r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        double za4, zb4;
        plom_normal_pair64(ns, &za4, &zb4);
        const double w0_8 = sqdt * za4;
        const double f8 = X[8] + ( - (r0_8*X[8]) + (r3_8))*dtr + - (r0_8*X[8])*P48*w0_8;
        const double f18 = X[18] + ( - (r1_8*X[18]) + (r0_8*X[8]))*dtr + + (r0_8*X[8])*P48*w0_8;
        const double f28 = X[28] + ( - (r2_8*X[28]) + (r1_8*X[18]))*dtr;
        s38 += ( + (r1_8*X[18]))*dtr;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        X[18] = (f18 < 0.0) ? 0.0 : f18;
        const double r0_9 = ((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])) / C9);
        const double r1_9 = P19;
        const double r2_9 = P29;
        const double r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        const double w0_9 = sqdt * zb4;
        const double f9 = X[9] + ( - (r0_9*X[9]) + (r3_9))*dtr + - (r0_9*X[9])*P49*w0_9;
        const double f19 = X[19] + ( - (r1_9*X[19]) + (r0_9*X[9]))*dtr + + (r0_9*X[9])*P49*w0_9;
        const double f29 = X[29] + ( - (r2_9*X[29]) + (r1_9*X[19]))*dtr;
        s39 += ( + (r1_9*X[19]))*dtr;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        X[19] = (f19 < 0.0) ? 0.0 : f19;
        X[20] = (f20 < 0.0) ? 0.0 : f20;
        X[21] = (f21 < 0.0) ? 0.0 : f21;
        X[22] = (f22 < 0.0) ? 0.0 : f22;
        X[23] = (f23 < 0.0) ? 0.0 : f23;
        X[24] = (f24 < 0.0) ? 0.0 : f24;
        X[25] = (f25 < 0.0) ? 0.0 : f25;
        X[26] = (f26 < 0.0) ? 0.0 : f26;
        X[27] = (f27 < 0.0) ? 0.0 : f27;
        X[28] = (f28 < 0.0) ? 0.0 : f28;
        X[29] = (f29 < 0.0) ? 0.0 : f29;
        const double f30 = X[30] + s30;
        X[30] = (f30 < 0.0) ? 0.0 : f30;
        const double f31 = X[31] + s31;
        X[31] = (f31 < 0.0) ? 0.0 : f31;
        const double f32 = X[32] + s32;
        X[32] = (f32 < 0.0) ? 0.0 : f32;
        const double f33 = X[33] + s33;
        X[33] = (f33 < 0.0) ? 0.0 : f33;
        const double f34 = X[34] + s34;
        X[34] = (f34 < 0.0) ? 0.0 : f34;
        const double f35 = X[35] + s35;
        X[35] = (f35 < 0.0) ? 0.0 : f35;
        const double f36 = X[36] + s36;
        X[36] = (f36 < 0.0) ? 0.0 : f36;
        const double f37 = X[37] + s37;
        X[37] = (f37 < 0.0) ? 0.0 : f37;
        const double f38 = X[38] + s38;
        X[38] = (f38 < 0.0) ? 0.0 : f38;
        const double f39 = X[39] + s39;
        X[39] = (f39 < 0.0) ? 0.0 : f39;
        t += dt;
    }
}

PLOM_HD void plom_seir10_interval_no_env_sto_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
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
    const double P20 = PLOM_UNIFORM64(pv[20]);
    const double P21 = PLOM_UNIFORM64(pv[21]);
    const double P22 = PLOM_UNIFORM64(pv[22]);
    const double P23 = PLOM_UNIFORM64(pv[23]);
    const double P24 = PLOM_UNIFORM64(pv[24]);
    const double P25 = PLOM_UNIFORM64(pv[25]);
    const double P26 = PLOM_UNIFORM64(pv[26]);
    const double P27 = PLOM_UNIFORM64(pv[27]);
    const double P28 = PLOM_UNIFORM64(pv[28]);
    const double P29 = PLOM_UNIFORM64(pv[29]);
    const double P30 = PLOM_UNIFORM64(pv[30]);
    const double P31 = PLOM_UNIFORM64(pv[31]);
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P80 = PLOM_UNIFORM64(pv[80]);
    const double P81 = PLOM_UNIFORM64(pv[81]);
    const double P82 = PLOM_UNIFORM64(pv[82]);
    const double P83 = PLOM_UNIFORM64(pv[83]);
    const double P84 = PLOM_UNIFORM64(pv[84]);
    const double P85 = PLOM_UNIFORM64(pv[85]);
    const double P86 = PLOM_UNIFORM64(pv[86]);
    const double P87 = PLOM_UNIFORM64(pv[87]);
    const double P88 = PLOM_UNIFORM64(pv[88]);
    const double P89 = PLOM_UNIFORM64(pv[89]);
    const double P90 = PLOM_UNIFORM64(pv[90]);
    const double P91 = PLOM_UNIFORM64(pv[91]);
    const double P92 = PLOM_UNIFORM64(pv[92]);
    const double P93 = PLOM_UNIFORM64(pv[93]);
    const double P94 = PLOM_UNIFORM64(pv[94]);
    const double P95 = PLOM_UNIFORM64(pv[95]);
    const double P96 = PLOM_UNIFORM64(pv[96]);
    const double P97 = PLOM_UNIFORM64(pv[97]);
    const double P98 = PLOM_UNIFORM64(pv[98]);
    const double P99 = PLOM_UNIFORM64(pv[99]);
    const double P100 = PLOM_UNIFORM64(pv[100]);
    const double P101 = PLOM_UNIFORM64(pv[101]);
    const double P102 = PLOM_UNIFORM64(pv[102]);
    const double P103 = PLOM_UNIFORM64(pv[103]);
    const double P104 = PLOM_UNIFORM64(pv[104]);
    const double P105 = PLOM_UNIFORM64(pv[105]);
    const double P106 = PLOM_UNIFORM64(pv[106]);
    const double P107 = PLOM_UNIFORM64(pv[107]);
    const double P108 = PLOM_UNIFORM64(pv[108]);
    const double P109 = PLOM_UNIFORM64(pv[109]);
    const double P110 = PLOM_UNIFORM64(pv[110]);
    const double P111 = PLOM_UNIFORM64(pv[111]);
    const double P112 = PLOM_UNIFORM64(pv[112]);
    const double P113 = PLOM_UNIFORM64(pv[113]);
    const double P114 = PLOM_UNIFORM64(pv[114]);
    const double P115 = PLOM_UNIFORM64(pv[115]);
    const double P116 = PLOM_UNIFORM64(pv[116]);
    const double P117 = PLOM_UNIFORM64(pv[117]);
    const double P118 = PLOM_UNIFORM64(pv[118]);
    const double P119 = PLOM_UNIFORM64(pv[119]);
    const double P120 = PLOM_UNIFORM64(pv[120]);
    const double P121 = PLOM_UNIFORM64(pv[121]);
    const double P122 = PLOM_UNIFORM64(pv[122]);
    const double P123 = PLOM_UNIFORM64(pv[123]);
    const double P124 = PLOM_UNIFORM64(pv[124]);
    const double P125 = PLOM_UNIFORM64(pv[125]);
    const double P126 = PLOM_UNIFORM64(pv[126]);
    const double P127 = PLOM_UNIFORM64(pv[127]);
    const double P128 = PLOM_UNIFORM64(pv[128]);
    const double P129 = PLOM_UNIFORM64(pv[129]);
    const double P130 = PLOM_UNIFORM64(pv[130]);
    const double P131 = PLOM_UNIFORM64(pv[131]);
    const double P132 = PLOM_UNIFORM64(pv[132]);
    const double P133 = PLOM_UNIFORM64(pv[133]);
    const double P134 = PLOM_UNIFORM64(pv[134]);
    const double P135 = PLOM_UNIFORM64(pv[135]);
    const double P136 = PLOM_UNIFORM64(pv[136]);
    const double P137 = PLOM_UNIFORM64(pv[137]);
    const double P138 = PLOM_UNIFORM64(pv[138]);
    const double P139 = PLOM_UNIFORM64(pv[139]);
    const double P140 = PLOM_UNIFORM64(pv[140]);
    const double P141 = PLOM_UNIFORM64(pv[141]);
    const double P142 = PLOM_UNIFORM64(pv[142]);
    const double P143 = PLOM_UNIFORM64(pv[143]);
    const double P144 = PLOM_UNIFORM64(pv[144]);
    const double P145 = PLOM_UNIFORM64(pv[145]);
    const double P146 = PLOM_UNIFORM64(pv[146]);
    const double P147 = PLOM_UNIFORM64(pv[147]);
    const double P148 = PLOM_UNIFORM64(pv[148]);
    const double P149 = PLOM_UNIFORM64(pv[149]);
    const double P150 = PLOM_UNIFORM64(pv[150]);
    const double P151 = PLOM_UNIFORM64(pv[151]);
    const double P152 = PLOM_UNIFORM64(pv[152]);
    const double P153 = PLOM_UNIFORM64(pv[153]);
    const double P154 = PLOM_UNIFORM64(pv[154]);
    const double P155 = PLOM_UNIFORM64(pv[155]);
    const double P156 = PLOM_UNIFORM64(pv[156]);
    const double P157 = PLOM_UNIFORM64(pv[157]);
    const double P158 = PLOM_UNIFORM64(pv[158]);
    const double P159 = PLOM_UNIFORM64(pv[159]);
    const double P160 = PLOM_UNIFORM64(pv[160]);
    const double P161 = PLOM_UNIFORM64(pv[161]);
    const double P162 = PLOM_UNIFORM64(pv[162]);
    const double P163 = PLOM_UNIFORM64(pv[163]);
    const double P164 = PLOM_UNIFORM64(pv[164]);
    const double P165 = PLOM_UNIFORM64(pv[165]);
    const double P166 = PLOM_UNIFORM64(pv[166]);
    const double P167 = PLOM_UNIFORM64(pv[167]);
    const double P168 = PLOM_UNIFORM64(pv[168]);
    const double P169 = PLOM_UNIFORM64(pv[169]);
    const double P170 = PLOM_UNIFORM64(pv[170]);
    const double P171 = PLOM_UNIFORM64(pv[171]);
    const double P172 = PLOM_UNIFORM64(pv[172]);
    const double P173 = PLOM_UNIFORM64(pv[173]);
    const double P174 = PLOM_UNIFORM64(pv[174]);
    const double P175 = PLOM_UNIFORM64(pv[175]);
    const double P176 = PLOM_UNIFORM64(pv[176]);
    const double P177 = PLOM_UNIFORM64(pv[177]);
    const double P178 = PLOM_UNIFORM64(pv[178]);
    const double P179 = PLOM_UNIFORM64(pv[179]);
    X[30] = 0.0; /* reset_inc */
    X[31] = 0.0; /* reset_inc */
    X[32] = 0.0; /* reset_inc */
    X[33] = 0.0; /* reset_inc */
    X[34] = 0.0; /* reset_inc */
    X[35] = 0.0; /* reset_inc */
    X[36] = 0.0; /* reset_inc */
    X[37] = 0.0; /* reset_inc */
    X[38] = 0.0; /* reset_inc */
    X[39] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double C4 = plom_cov_ld(cvk + 4);
        const double C5 = plom_cov_ld(cvk + 5);
        const double C6 = plom_cov_ld(cvk + 6);
        const double C7 = plom_cov_ld(cvk + 7);
        const double C8 = plom_cov_ld(cvk + 8);
        const double C9 = plom_cov_ld(cvk + 9);
        double s30 = 0.0;
        double s31 = 0.0;
        double s32 = 0.0;
        double s33 = 0.0;
        double s34 = 0.0;
        double s35 = 0.0;
        double s36 = 0.0;
        double s37 = 0.0;
        double s38 = 0.0;
        double s39 = 0.0;
        const double r0_0 = ((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])) / C0);
        const double r1_0 = P10;
        const double r2_0 = P20;
        const double r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        double za0, zb0;
        plom_normal_pair64(ns, &za0, &zb0);
        const double d0_0 = sqdt * za0;
        const double d1_0 = sqdt * zb0;
        double za1, zb1;
        plom_normal_pair64(ns, &za1, &zb1);
        const double d2_0 = sqdt * za1;
        const double d3_0 = sqdt * zb1;
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r3_0))*dtr + - plom_sqrt_rate((r0_0*X[0]))*d0_0+ plom_sqrt_rate((r3_0))*d3_0;
        const double f10 = X[10] + ( - (r1_0*X[10]) + (r0_0*X[0]))*dtr + - plom_sqrt_rate((r1_0*X[10]))*d1_0+ plom_sqrt_rate((r0_0*X[0]))*d0_0;
        const double f20 = X[20] + ( - (r2_0*X[20]) + (r1_0*X[10]))*dtr + - plom_sqrt_rate((r2_0*X[20]))*d2_0+ plom_sqrt_rate((r1_0*X[10]))*d1_0;
        s30 += ( + (r1_0*X[10]))*dtr + + plom_sqrt_rate((r1_0*X[10]))*d1_0;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        const double r0_1 = ((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])) / C1);
        const double r1_1 = P11;
        const double r2_1 = P21;
        const double r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        double za2, zb2;
        plom_normal_pair64(ns, &za2, &zb2);
        const double d0_1 = sqdt * za2;
        const double d1_1 = sqdt * zb2;
        double za3, zb3;
        plom_normal_pair64(ns, &za3, &zb3);
        const double d2_1 = sqdt * za3;
        const double d3_1 = sqdt * zb3;
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r3_1))*dtr + - plom_sqrt_rate((r0_1*X[1]))*d0_1+ plom_sqrt_rate((r3_1))*d3_1;
        const double f11 = X[11] + ( - (r1_1*X[11]) + (r0_1*X[1]))*dtr + - plom_sqrt_rate((r1_1*X[11]))*d1_1+ plom_sqrt_rate((r0_1*X[1]))*d0_1;
        const double f21 = X[21] + ( - (r2_1*X[21]) + (r1_1*X[11]))*dtr + - plom_sqrt_rate((r2_1*X[21]))*d2_1+ plom_sqrt_rate((r1_1*X[11]))*d1_1;
        s31 += ( + (r1_1*X[11]))*dtr + + plom_sqrt_rate((r1_1*X[11]))*d1_1;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        const double r0_2 = ((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])) / C2);
        const double r1_2 = P12;
        const double r2_2 = P22;
        const double r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        double za4, zb4;
        plom_normal_pair64(ns, &za4, &zb4);
        const double d0_2 = sqdt * za4;
        const double d1_2 = sqdt * zb4;
        double za5, zb5;
        plom_normal_pair64(ns, &za5, &zb5);
        const double d2_2 = sqdt * za5;
        const double d3_2 = sqdt * zb5;
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r3_2))*dtr + - plom_sqrt_rate((r0_2*X[2]))*d0_2+ plom_sqrt_rate((r3_2))*d3_2;
        const double f12 = X[12] + ( - (r1_2*X[12]) + (r0_2*X[2]))*dtr + - plom_sqrt_rate((r1_2*X[12]))*d1_2+ plom_sqrt_rate((r0_2*X[2]))*d0_2;
        const double f22 = X[22] + ( - (r2_2*X[22]) + (r1_2*X[12]))*dtr + - plom_sqrt_rate((r2_2*X[22]))*d2_2+ plom_sqrt_rate((r1_2*X[12]))*d1_2;
        s32 += ( + (r1_2*X[12]))*dtr + + plom_sqrt_rate((r1_2*X[12]))*d1_2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        X[12] = (f12 < 0.0) ? 0.0 : f12;
        const double r0_3 = ((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])) / C3);
        const double r1_3 = P13;
        const double r2_3 = P23;
        const double r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        double za6, zb6;
        plom_normal_pair64(ns, &za6, &zb6);
        const double d0_3 = sqdt * za6;
        const double d1_3 = sqdt * zb6;
        double za7, zb7;
        plom_normal_pair64(ns, &za7, &zb7);
        const double d2_3 = sqdt * za7;
        const double d3_3 = sqdt * zb7;
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r3_3))*dtr + - plom_sqrt_rate((r0_3*X[3]))*d0_3+ plom_sqrt_rate((r3_3))*d3_3;
        const double f13 = X[13] + ( - (r1_3*X[13]) + (r0_3*X[3]))*dtr + - plom_sqrt_rate((r1_3*X[13]))*d1_3+ plom_sqrt_rate((r0_3*X[3]))*d0_3;
        const double f23 = X[23] + ( - (r2_3*X[23]) + (r1_3*X[13]))*dtr + - plom_sqrt_rate((r2_3*X[23]))*d2_3+ plom_sqrt_rate((r1_3*X[13]))*d1_3;
        s33 += ( + (r1_3*X[13]))*dtr + + plom_sqrt_rate((r1_3*X[13]))*d1_3;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[13] = (f13 < 0.0) ? 0.0 : f13;
        const double r0_4 = ((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])) / C4);
        const double r1_4 = P14;
        const double r2_4 = P24;
        const double r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        double za8, zb8;
        plom_normal_pair64(ns, &za8, &zb8);
        const double d0_4 = sqdt * za8;
        const double d1_4 = sqdt * zb8;
        double za9, zb9;
        plom_normal_pair64(ns, &za9, &zb9);
        const double d2_4 = sqdt * za9;
        const double d3_4 = sqdt * zb9;
        const double f4 = X[4] + ( - (r0_4*X[4]) + (r3_4))*dtr + - plom_sqrt_rate((r0_4*X[4]))*d0_4+ plom_sqrt_rate((r3_4))*d3_4;
        const double f14 = X[14] + ( - (r1_4*X[14]) + (r0_4*X[4]))*dtr + - plom_sqrt_rate((r1_4*X[14]))*d1_4+ plom_sqrt_rate((r0_4*X[4]))*d0_4;
        const double f24 = X[24] + ( - (r2_4*X[24]) + (r1_4*X[14]))*dtr + - plom_sqrt_rate((r2_4*X[24]))*d2_4+ plom_sqrt_rate((r1_4*X[14]))*d1_4;
        s34 += ( + (r1_4*X[14]))*dtr + + plom_sqrt_rate((r1_4*X[14]))*d1_4;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[14] = (f14 < 0.0) ? 0.0 : f14;
        const double r0_5 = ((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])) / C5);
        const double r1_5 = P15;
        const double r2_5 = P25;
        const double r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        double za10, zb10;
        plom_normal_pair64(ns, &za10, &zb10);
        const double d0_5 = sqdt * za10;
        const double d1_5 = sqdt * zb10;
        double za11, zb11;
        plom_normal_pair64(ns, &za11, &zb11);
        const double d2_5 = sqdt * za11;
        const double d3_5 = sqdt * zb11;
        const double f5 = X[5] + ( - (r0_5*X[5]) + (r3_5))*dtr + - plom_sqrt_rate((r0_5*X[5]))*d0_5+ plom_sqrt_rate((r3_5))*d3_5;
        const double f15 = X[15] + ( - (r1_5*X[15]) + (r0_5*X[5]))*dtr + - plom_sqrt_rate((r1_5*X[15]))*d1_5+ plom_sqrt_rate((r0_5*X[5]))*d0_5;
        const double f25 = X[25] + ( - (r2_5*X[25]) + (r1_5*X[15]))*dtr + - plom_sqrt_rate((r2_5*X[25]))*d2_5+ plom_sqrt_rate((r1_5*X[15]))*d1_5;
        s35 += ( + (r1_5*X[15]))*dtr + + plom_sqrt_rate((r1_5*X[15]))*d1_5;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[15] = (f15 < 0.0) ? 0.0 : f15;
        const double r0_6 = ((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])) / C6);
        const double r1_6 = P16;
        const double r2_6 = P26;
        const double r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        double za12, zb12;
        plom_normal_pair64(ns, &za12, &zb12);
        const double d0_6 = sqdt * za12;
        const double d1_6 = sqdt * zb12;
        double za13, zb13;
        plom_normal_pair64(ns, &za13, &zb13);
        const double d2_6 = sqdt * za13;
        const double d3_6 = sqdt * zb13;
        const double f6 = X[6] + ( - (r0_6*X[6]) + (r3_6))*dtr + - plom_sqrt_rate((r0_6*X[6]))*d0_6+ plom_sqrt_rate((r3_6))*d3_6;
        const double f16 = X[16] + ( - (r1_6*X[16]) + (r0_6*X[6]))*dtr + - plom_sqrt_rate((r1_6*X[16]))*d1_6+ plom_sqrt_rate((r0_6*X[6]))*d0_6;
        const double f26 = X[26] + ( - (r2_6*X[26]) + (r1_6*X[16]))*dtr + - plom_sqrt_rate((r2_6*X[26]))*d2_6+ plom_sqrt_rate((r1_6*X[16]))*d1_6;
        s36 += ( + (r1_6*X[16]))*dtr + + plom_sqrt_rate((r1_6*X[16]))*d1_6;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[16] = (f16 < 0.0) ? 0.0 : f16;
        const double r0_7 = ((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])) / C7);
        const double r1_7 = P17;
        const double r2_7 = P27;
        const double r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        double za14, zb14;
        plom_normal_pair64(ns, &za14, &zb14);
        const double d0_7 = sqdt * za14;
        const double d1_7 = sqdt * zb14;
        double za15, zb15;
        plom_normal_pair64(ns, &za15, &zb15);
        const double d2_7 = sqdt * za15;
        const double d3_7 = sqdt * zb15;
        const double f7 = X[7] + ( - (r0_7*X[7]) + (r3_7))*dtr + - plom_sqrt_rate((r0_7*X[7]))*d0_7+ plom_sqrt_rate((r3_7))*d3_7;
        const double f17 = X[17] + ( - (r1_7*X[17]) + (r0_7*X[7]))*dtr + - plom_sqrt_rate((r1_7*X[17]))*d1_7+ plom_sqrt_rate((r0_7*X[7]))*d0_7;
        const double f27 = X[27] + ( - (r2_7*X[27]) + (r1_7*X[17]))*dtr + - plom_sqrt_rate((r2_7*X[27]))*d2_7+ plom_sqrt_rate((r1_7*X[17]))*d1_7;
        s37 += ( + (r1_7*X[17]))*dtr + + plom_sqrt_rate((r1_7*X[17]))*d1_7;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        X[17] = (f17 < 0.0) ? 0.0 : f17;
        const double r0_8 = ((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])) / C8);
        const double r1_8 = P18;
        const double r2_8 = P28;
        const double r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        double za16, zb16;
        plom_normal_pair64(ns, &za16, &zb16);
        const double d0_8 = sqdt * za16;
        const double d1_8 = sqdt * zb16;
        double za17, zb17;
        plom_normal_pair64(ns, &za17, &zb17);
        const double d2_8 = sqdt * za17;
        const double d3_8 = sqdt * zb17;
        const double f8 = X[8] + ( - (r0_8*X[8]) + (r3_8))*dtr + - plom_sqrt_rate((r0_8*X[8]))*d0_8+ plom_sqrt_rate((r3_8))*d3_8;
        const double f18 = X[18] + ( - (r1_8*X[18]) + (r0_8*X[8]))*dtr + - plom_sqrt_rate((r1_8*X[18]))*d1_8+ plom_sqrt_rate((r0_8*X[8]))*d0_8;
        const double f28 = X[28] + ( - (r2_8*X[28]) + (r1_8*X[18]))*dtr + - plom_sqrt_rate((r2_8*X[28]))*d2_8+ plom_sqrt_rate((r1_8*X[18]))*d1_8;
        s38 += ( + (r1_8*X[18]))*dtr + + plom_sqrt_rate((r1_8*X[18]))*d1_8;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        X[18] = (f18 < 0.0) ? 0.0 : f18;
        const double r0_9 = ((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])) / C9);
        const double r1_9 = P19;
        const double r2_9 = P29;
        const double r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        double za18, zb18;
        plom_normal_pair64(ns, &za18, &zb18);
        const double d0_9 = sqdt * za18;
        const double d1_9 = sqdt * zb18;
        double za19, zb19;
        plom_normal_pair64(ns, &za19, &zb19);
        const double d2_9 = sqdt * za19;
        const double d3_9 = sqdt * zb19;
        const double f9 = X[9] + ( - (r0_9*X[9]) + (r3_9))*dtr + - plom_sqrt_rate((r0_9*X[9]))*d0_9+ plom_sqrt_rate((r3_9))*d3_9;
        const double f19 = X[19] + ( - (r1_9*X[19]) + (r0_9*X[9]))*dtr + - plom_sqrt_rate((r1_9*X[19]))*d1_9+ plom_sqrt_rate((r0_9*X[9]))*d0_9;
        const double f29 = X[29] + ( - (r2_9*X[29]) + (r1_9*X[19]))*dtr + - plom_sqrt_rate((r2_9*X[29]))*d2_9+ plom_sqrt_rate((r1_9*X[19]))*d1_9;
        s39 += ( + (r1_9*X[19]))*dtr + + plom_sqrt_rate((r1_9*X[19]))*d1_9;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        X[19] = (f19 < 0.0) ? 0.0 : f19;
        X[20] = (f20 < 0.0) ? 0.0 : f20;
        X[21] = (f21 < 0.0) ? 0.0 : f21;
        X[22] = (f22 < 0.0) ? 0.0 : f22;
        X[23] = (f23 < 0.0) ? 0.0 : f23;
        X[24] = (f24 < 0.0) ? 0.0 : f24;
        X[25] = (f25 < 0.0) ? 0.0 : f25;
        X[26] = (f26 < 0.0) ? 0.0 : f26;
        X[27] = (f27 < 0.0) ? 0.0 : f27;
        X[28] = (f28 < 0.0) ? 0.0 : f28;
        X[29] = (f29 < 0.0) ? 0.0 : f29;
        const double f30 = X[30] + s30;
        X[30] = (f30 < 0.0) ? 0.0 : f30;
        const double f31 = X[31] + s31;
        X[31] = (f31 < 0.0) ? 0.0 : f31;
        const double f32 = X[32] + s32;
        X[32] = (f32 < 0.0) ? 0.0 : f32;
        const double f33 = X[33] + s33;
        X[33] = (f33 < 0.0) ? 0.0 : f33;
        const double f34 = X[34] + s34;
        X[34] = (f34 < 0.0) ? 0.0 : f34;
        const double f35 = X[35] + s35;
        X[35] = (f35 < 0.0) ? 0.0 : f35;
        const double f36 = X[36] + s36;
        X[36] = (f36 < 0.0) ? 0.0 : f36;
        const double f37 = X[37] + s37;
        X[37] = (f37 < 0.0) ? 0.0 : f37;
        const double f38 = X[38] + s38;
        X[38] = (f38 < 0.0) ? 0.0 : f38;
        const double f39 = X[39] + s39;
        X[39] = (f39 < 0.0) ? 0.0 : f39;
        t += dt;
    }
}

PLOM_HD void plom_seir10_interval_no_dem_sto_no_env_sto_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
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
    const double P20 = PLOM_UNIFORM64(pv[20]);
    const double P21 = PLOM_UNIFORM64(pv[21]);
    const double P22 = PLOM_UNIFORM64(pv[22]);
    const double P23 = PLOM_UNIFORM64(pv[23]);
    const double P24 = PLOM_UNIFORM64(pv[24]);
    const double P25 = PLOM_UNIFORM64(pv[25]);
    const double P26 = PLOM_UNIFORM64(pv[26]);
    const double P27 = PLOM_UNIFORM64(pv[27]);
    const double P28 = PLOM_UNIFORM64(pv[28]);
    const double P29 = PLOM_UNIFORM64(pv[29]);
    const double P30 = PLOM_UNIFORM64(pv[30]);
    const double P31 = PLOM_UNIFORM64(pv[31]);
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P80 = PLOM_UNIFORM64(pv[80]);
    const double P81 = PLOM_UNIFORM64(pv[81]);
    const double P82 = PLOM_UNIFORM64(pv[82]);
    const double P83 = PLOM_UNIFORM64(pv[83]);
    const double P84 = PLOM_UNIFORM64(pv[84]);
    const double P85 = PLOM_UNIFORM64(pv[85]);
    const double P86 = PLOM_UNIFORM64(pv[86]);
    const double P87 = PLOM_UNIFORM64(pv[87]);
    const double P88 = PLOM_UNIFORM64(pv[88]);
    const double P89 = PLOM_UNIFORM64(pv[89]);
    const double P90 = PLOM_UNIFORM64(pv[90]);
    const double P91 = PLOM_UNIFORM64(pv[91]);
    const double P92 = PLOM_UNIFORM64(pv[92]);
    const double P93 = PLOM_UNIFORM64(pv[93]);
    const double P94 = PLOM_UNIFORM64(pv[94]);
    const double P95 = PLOM_UNIFORM64(pv[95]);
    const double P96 = PLOM_UNIFORM64(pv[96]);
    const double P97 = PLOM_UNIFORM64(pv[97]);
    const double P98 = PLOM_UNIFORM64(pv[98]);
    const double P99 = PLOM_UNIFORM64(pv[99]);
    const double P100 = PLOM_UNIFORM64(pv[100]);
    const double P101 = PLOM_UNIFORM64(pv[101]);
    const double P102 = PLOM_UNIFORM64(pv[102]);
    const double P103 = PLOM_UNIFORM64(pv[103]);
    const double P104 = PLOM_UNIFORM64(pv[104]);
    const double P105 = PLOM_UNIFORM64(pv[105]);
    const double P106 = PLOM_UNIFORM64(pv[106]);
    const double P107 = PLOM_UNIFORM64(pv[107]);
    const double P108 = PLOM_UNIFORM64(pv[108]);
    const double P109 = PLOM_UNIFORM64(pv[109]);
    const double P110 = PLOM_UNIFORM64(pv[110]);
    const double P111 = PLOM_UNIFORM64(pv[111]);
    const double P112 = PLOM_UNIFORM64(pv[112]);
    const double P113 = PLOM_UNIFORM64(pv[113]);
    const double P114 = PLOM_UNIFORM64(pv[114]);
    const double P115 = PLOM_UNIFORM64(pv[115]);
    const double P116 = PLOM_UNIFORM64(pv[116]);
    const double P117 = PLOM_UNIFORM64(pv[117]);
    const double P118 = PLOM_UNIFORM64(pv[118]);
    const double P119 = PLOM_UNIFORM64(pv[119]);
    const double P120 = PLOM_UNIFORM64(pv[120]);
    const double P121 = PLOM_UNIFORM64(pv[121]);
    const double P122 = PLOM_UNIFORM64(pv[122]);
    const double P123 = PLOM_UNIFORM64(pv[123]);
    const double P124 = PLOM_UNIFORM64(pv[124]);
    const double P125 = PLOM_UNIFORM64(pv[125]);
    const double P126 = PLOM_UNIFORM64(pv[126]);
    const double P127 = PLOM_UNIFORM64(pv[127]);
    const double P128 = PLOM_UNIFORM64(pv[128]);
    const double P129 = PLOM_UNIFORM64(pv[129]);
    const double P130 = PLOM_UNIFORM64(pv[130]);
    const double P131 = PLOM_UNIFORM64(pv[131]);
    const double P132 = PLOM_UNIFORM64(pv[132]);
    const double P133 = PLOM_UNIFORM64(pv[133]);
    const double P134 = PLOM_UNIFORM64(pv[134]);
    const double P135 = PLOM_UNIFORM64(pv[135]);
    const double P136 = PLOM_UNIFORM64(pv[136]);
    const double P137 = PLOM_UNIFORM64(pv[137]);
    const double P138 = PLOM_UNIFORM64(pv[138]);
    const double P139 = PLOM_UNIFORM64(pv[139]);
    const double P140 = PLOM_UNIFORM64(pv[140]);
    const double P141 = PLOM_UNIFORM64(pv[141]);
    const double P142 = PLOM_UNIFORM64(pv[142]);
    const double P143 = PLOM_UNIFORM64(pv[143]);
    const double P144 = PLOM_UNIFORM64(pv[144]);
    const double P145 = PLOM_UNIFORM64(pv[145]);
    const double P146 = PLOM_UNIFORM64(pv[146]);
    const double P147 = PLOM_UNIFORM64(pv[147]);
    const double P148 = PLOM_UNIFORM64(pv[148]);
    const double P149 = PLOM_UNIFORM64(pv[149]);
    const double P150 = PLOM_UNIFORM64(pv[150]);
    const double P151 = PLOM_UNIFORM64(pv[151]);
    const double P152 = PLOM_UNIFORM64(pv[152]);
    const double P153 = PLOM_UNIFORM64(pv[153]);
    const double P154 = PLOM_UNIFORM64(pv[154]);
    const double P155 = PLOM_UNIFORM64(pv[155]);
    const double P156 = PLOM_UNIFORM64(pv[156]);
    const double P157 = PLOM_UNIFORM64(pv[157]);
    const double P158 = PLOM_UNIFORM64(pv[158]);
    const double P159 = PLOM_UNIFORM64(pv[159]);
    const double P160 = PLOM_UNIFORM64(pv[160]);
    const double P161 = PLOM_UNIFORM64(pv[161]);
    const double P162 = PLOM_UNIFORM64(pv[162]);
    const double P163 = PLOM_UNIFORM64(pv[163]);
    const double P164 = PLOM_UNIFORM64(pv[164]);
    const double P165 = PLOM_UNIFORM64(pv[165]);
    const double P166 = PLOM_UNIFORM64(pv[166]);
    const double P167 = PLOM_UNIFORM64(pv[167]);
    const double P168 = PLOM_UNIFORM64(pv[168]);
    const double P169 = PLOM_UNIFORM64(pv[169]);
    const double P170 = PLOM_UNIFORM64(pv[170]);
    const double P171 = PLOM_UNIFORM64(pv[171]);
    const double P172 = PLOM_UNIFORM64(pv[172]);
    const double P173 = PLOM_UNIFORM64(pv[173]);
    const double P174 = PLOM_UNIFORM64(pv[174]);
    const double P175 = PLOM_UNIFORM64(pv[175]);
    const double P176 = PLOM_UNIFORM64(pv[176]);
    const double P177 = PLOM_UNIFORM64(pv[177]);
    const double P178 = PLOM_UNIFORM64(pv[178]);
    const double P179 = PLOM_UNIFORM64(pv[179]);
    X[30] = 0.0; /* reset_inc */
    X[31] = 0.0; /* reset_inc */
    X[32] = 0.0; /* reset_inc */
    X[33] = 0.0; /* reset_inc */
    X[34] = 0.0; /* reset_inc */
    X[35] = 0.0; /* reset_inc */
    X[36] = 0.0; /* reset_inc */
    X[37] = 0.0; /* reset_inc */
    X[38] = 0.0; /* reset_inc */
    X[39] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double C4 = plom_cov_ld(cvk + 4);
        const double C5 = plom_cov_ld(cvk + 5);
        const double C6 = plom_cov_ld(cvk + 6);
        const double C7 = plom_cov_ld(cvk + 7);
        const double C8 = plom_cov_ld(cvk + 8);
        const double C9 = plom_cov_ld(cvk + 9);
        double s30 = 0.0;
        double s31 = 0.0;
        double s32 = 0.0;
        double s33 = 0.0;
        double s34 = 0.0;
        double s35 = 0.0;
        double s36 = 0.0;
        double s37 = 0.0;
        double s38 = 0.0;
        double s39 = 0.0;
        const double r0_0 = ((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])) / C0);
        const double r1_0 = P10;
        const double r2_0 = P20;
        const double r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r3_0))*dtr;
        const double f10 = X[10] + ( - (r1_0*X[10]) + (r0_0*X[0]))*dtr;
        const double f20 = X[20] + ( - (r2_0*X[20]) + (r1_0*X[10]))*dtr;
        s30 += ( + (r1_0*X[10]))*dtr;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        const double r0_1 = ((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])) / C1);
        const double r1_1 = P11;
        const double r2_1 = P21;
        const double r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r3_1))*dtr;
        const double f11 = X[11] + ( - (r1_1*X[11]) + (r0_1*X[1]))*dtr;
        const double f21 = X[21] + ( - (r2_1*X[21]) + (r1_1*X[11]))*dtr;
        s31 += ( + (r1_1*X[11]))*dtr;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        const double r0_2 = ((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])) / C2);
        const double r1_2 = P12;
        const double r2_2 = P22;
        const double r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r3_2))*dtr;
        const double f12 = X[12] + ( - (r1_2*X[12]) + (r0_2*X[2]))*dtr;
        const double f22 = X[22] + ( - (r2_2*X[22]) + (r1_2*X[12]))*dtr;
        s32 += ( + (r1_2*X[12]))*dtr;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        X[12] = (f12 < 0.0) ? 0.0 : f12;
        const double r0_3 = ((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])) / C3);
        const double r1_3 = P13;
        const double r2_3 = P23;
        const double r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r3_3))*dtr;
        const double f13 = X[13] + ( - (r1_3*X[13]) + (r0_3*X[3]))*dtr;
        const double f23 = X[23] + ( - (r2_3*X[23]) + (r1_3*X[13]))*dtr;
        s33 += ( + (r1_3*X[13]))*dtr;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[13] = (f13 < 0.0) ? 0.0 : f13;
        const double r0_4 = ((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])) / C4);
        const double r1_4 = P14;
        const double r2_4 = P24;
        const double r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        const double f4 = X[4] + ( - (r0_4*X[4]) + (r3_4))*dtr;
        const double f14 = X[14] + ( - (r1_4*X[14]) + (r0_4*X[4]))*dtr;
        const double f24 = X[24] + ( - (r2_4*X[24]) + (r1_4*X[14]))*dtr;
        s34 += ( + (r1_4*X[14]))*dtr;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[14] = (f14 < 0.0) ? 0.0 : f14;
        const double r0_5 = ((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])) / C5);
        const double r1_5 = P15;
        const double r2_5 = P25;
        const double r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        const double f5 = X[5] + ( - (r0_5*X[5]) + (r3_5))*dtr;
        const double f15 = X[15] + ( - (r1_5*X[15]) + (r0_5*X[5]))*dtr;
        const double f25 = X[25] + ( - (r2_5*X[25]) + (r1_5*X[15]))*dtr;
        s35 += ( + (r1_5*X[15]))*dtr;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[15] = (f15 < 0.0) ? 0.0 : f15;
        const double r0_6 = ((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])) / C6);
        const double r1_6 = P16;
        const double r2_6 = P26;
        const double r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        const double f6 = X[6] + ( - (r0_6*X[6]) + (r3_6))*dtr;
        const double f16 = X[16] + ( - (r1_6*X[16]) + (r0_6*X[6]))*dtr;
        const double f26 = X[26] + ( - (r2_6*X[26]) + (r1_6*X[16]))*dtr;
        s36 += ( + (r1_6*X[16]))*dtr;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[16] = (f16 < 0.0) ? 0.0 : f16;
        const double r0_7 = ((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])) / C7);
        const double r1_7 = P17;
        const double r2_7 = P27;
        const double r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        const double f7 = X[7] + ( - (r0_7*X[7]) + (r3_7))*dtr;
        const double f17 = X[17] + ( - (r1_7*X[17]) + (r0_7*X[7]))*dtr;
        const double f27 = X[27] + ( - (r2_7*X[27]) + (r1_7*X[17]))*dtr;
        s37 += ( + (r1_7*X[17]))*dtr;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        X[17] = (f17 < 0.0) ? 0.0 : f17;
        const double r0_8 = ((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])) / C8);
        const double r1_8 = P18;
        const double r2_8 = P28;
        const double r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        const double f8 = X[8] + ( - (r0_8*X[8]) + (r3_8))*dtr;
        const double f18 = X[18] + ( - (r1_8*X[18]) + (r0_8*X[8]))*dtr;
        const double f28 = X[28] + ( - (r2_8*X[28]) + (r1_8*X[18]))*dtr;
        s38 += ( + (r1_8*X[18]))*dtr;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        X[18] = (f18 < 0.0) ? 0.0 : f18;
        const double r0_9 = ((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])) / C9);
        const double r1_9 = P19;
        const double r2_9 = P29;
        const double r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        const double f9 = X[9] + ( - (r0_9*X[9]) + (r3_9))*dtr;
        const double f19 = X[19] + ( - (r1_9*X[19]) + (r0_9*X[9]))*dtr;
        const double f29 = X[29] + ( - (r2_9*X[29]) + (r1_9*X[19]))*dtr;
        s39 += ( + (r1_9*X[19]))*dtr;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        X[19] = (f19 < 0.0) ? 0.0 : f19;
        X[20] = (f20 < 0.0) ? 0.0 : f20;
        X[21] = (f21 < 0.0) ? 0.0 : f21;
        X[22] = (f22 < 0.0) ? 0.0 : f22;
        X[23] = (f23 < 0.0) ? 0.0 : f23;
        X[24] = (f24 < 0.0) ? 0.0 : f24;
        X[25] = (f25 < 0.0) ? 0.0 : f25;
        X[26] = (f26 < 0.0) ? 0.0 : f26;
        X[27] = (f27 < 0.0) ? 0.0 : f27;
        X[28] = (f28 < 0.0) ? 0.0 : f28;
        X[29] = (f29 < 0.0) ? 0.0 : f29;
        const double f30 = X[30] + s30;
        X[30] = (f30 < 0.0) ? 0.0 : f30;
        const double f31 = X[31] + s31;
        X[31] = (f31 < 0.0) ? 0.0 : f31;
        const double f32 = X[32] + s32;
        X[32] = (f32 < 0.0) ? 0.0 : f32;
        const double f33 = X[33] + s33;
        X[33] = (f33 < 0.0) ? 0.0 : f33;
        const double f34 = X[34] + s34;
        X[34] = (f34 < 0.0) ? 0.0 : f34;
        const double f35 = X[35] + s35;
        X[35] = (f35 < 0.0) ? 0.0 : f35;
        const double f36 = X[36] + s36;
        X[36] = (f36 < 0.0) ? 0.0 : f36;
        const double f37 = X[37] + s37;
        X[37] = (f37 < 0.0) ? 0.0 : f37;
        const double f38 = X[38] + s38;
        X[38] = (f38 < 0.0) ? 0.0 : f38;
        const double f39 = X[39] + s39;
        X[39] = (f39 < 0.0) ? 0.0 : f39;
        t += dt;
    }
}

PLOM_HD void plom_seir10_proj2obs_f64(const double *X, const double *pvd, const double *cvd, double *obs)
{
    (void)pvd; (void)cvd; (void)X;
    obs[0] = (double)X[30];
    obs[1] = (double)X[31];
    obs[2] = (double)X[32];
    obs[3] = (double)X[33];
    obs[4] = (double)X[34];
    obs[5] = (double)X[35];
    obs[6] = (double)X[36];
    obs[7] = (double)X[37];
    obs[8] = (double)X[38];
    obs[9] = (double)X[39];
}

PLOM_HD double plom_seir10_loglik(const double *obs, const double *y, const double *pvd, const double *cvd, double t, double like_min, int *n_floor)
{
    double loglike = 0.0;
    int nf = 0;
    (void)t; (void)pvd; (void)cvd;
    if (!plom_isnan(y[0])) {
        const double x = obs[0];
        const double gsl_mu = (pvd[50] * x);
        const double gsl_sd = plom_sqrt((((pvd[50] * (1.0 - pvd[50])) * x) + plom_powi(((pvd[50] * pvd[60]) * x), 2)));
        double like;
        if (y[0] > 0.0) like = plom_cdf_gaussian_P(y[0] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[0] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[0] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[1])) {
        const double x = obs[1];
        const double gsl_mu = (pvd[51] * x);
        const double gsl_sd = plom_sqrt((((pvd[51] * (1.0 - pvd[51])) * x) + plom_powi(((pvd[51] * pvd[61]) * x), 2)));
        double like;
        if (y[1] > 0.0) like = plom_cdf_gaussian_P(y[1] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[1] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[1] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[2])) {
        const double x = obs[2];
        const double gsl_mu = (pvd[52] * x);
        const double gsl_sd = plom_sqrt((((pvd[52] * (1.0 - pvd[52])) * x) + plom_powi(((pvd[52] * pvd[62]) * x), 2)));
        double like;
        if (y[2] > 0.0) like = plom_cdf_gaussian_P(y[2] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[2] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[2] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[3])) {
        const double x = obs[3];
        const double gsl_mu = (pvd[53] * x);
        const double gsl_sd = plom_sqrt((((pvd[53] * (1.0 - pvd[53])) * x) + plom_powi(((pvd[53] * pvd[63]) * x), 2)));
        double like;
        if (y[3] > 0.0) like = plom_cdf_gaussian_P(y[3] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[3] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[3] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[4])) {
        const double x = obs[4];
        const double gsl_mu = (pvd[54] * x);
        const double gsl_sd = plom_sqrt((((pvd[54] * (1.0 - pvd[54])) * x) + plom_powi(((pvd[54] * pvd[64]) * x), 2)));
        double like;
        if (y[4] > 0.0) like = plom_cdf_gaussian_P(y[4] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[4] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[4] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[5])) {
        const double x = obs[5];
        const double gsl_mu = (pvd[55] * x);
        const double gsl_sd = plom_sqrt((((pvd[55] * (1.0 - pvd[55])) * x) + plom_powi(((pvd[55] * pvd[65]) * x), 2)));
        double like;
        if (y[5] > 0.0) like = plom_cdf_gaussian_P(y[5] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[5] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[5] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[6])) {
        const double x = obs[6];
        const double gsl_mu = (pvd[56] * x);
        const double gsl_sd = plom_sqrt((((pvd[56] * (1.0 - pvd[56])) * x) + plom_powi(((pvd[56] * pvd[66]) * x), 2)));
        double like;
        if (y[6] > 0.0) like = plom_cdf_gaussian_P(y[6] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[6] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[6] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[7])) {
        const double x = obs[7];
        const double gsl_mu = (pvd[57] * x);
        const double gsl_sd = plom_sqrt((((pvd[57] * (1.0 - pvd[57])) * x) + plom_powi(((pvd[57] * pvd[67]) * x), 2)));
        double like;
        if (y[7] > 0.0) like = plom_cdf_gaussian_P(y[7] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[7] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[7] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[8])) {
        const double x = obs[8];
        const double gsl_mu = (pvd[58] * x);
        const double gsl_sd = plom_sqrt((((pvd[58] * (1.0 - pvd[58])) * x) + plom_powi(((pvd[58] * pvd[68]) * x), 2)));
        double like;
        if (y[8] > 0.0) like = plom_cdf_gaussian_P(y[8] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[8] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[8] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[9])) {
        const double x = obs[9];
        const double gsl_mu = (pvd[59] * x);
        const double gsl_sd = plom_sqrt((((pvd[59] * (1.0 - pvd[59])) * x) + plom_powi(((pvd[59] * pvd[69]) * x), 2)));
        double like;
        if (y[9] > 0.0) like = plom_cdf_gaussian_P(y[9] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[9] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[9] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    *n_floor = nf;
    return loglike;
}

PLOM_HD void plom_seir10_obs_moments(const double *obs, const double *pvd, const double *cvd, double t, double *mean, double *var)
{
    (void)t; (void)pvd; (void)cvd;
    {
        const double x = obs[0];
        mean[0] = (pvd[50] * x);
        var[0] = (((pvd[50] * (1.0 - pvd[50])) * x) + plom_powi(((pvd[50] * pvd[60]) * x), 2));
    }
    {
        const double x = obs[1];
        mean[1] = (pvd[51] * x);
        var[1] = (((pvd[51] * (1.0 - pvd[51])) * x) + plom_powi(((pvd[51] * pvd[61]) * x), 2));
    }
    {
        const double x = obs[2];
        mean[2] = (pvd[52] * x);
        var[2] = (((pvd[52] * (1.0 - pvd[52])) * x) + plom_powi(((pvd[52] * pvd[62]) * x), 2));
    }
    {
        const double x = obs[3];
        mean[3] = (pvd[53] * x);
        var[3] = (((pvd[53] * (1.0 - pvd[53])) * x) + plom_powi(((pvd[53] * pvd[63]) * x), 2));
    }
    {
        const double x = obs[4];
        mean[4] = (pvd[54] * x);
        var[4] = (((pvd[54] * (1.0 - pvd[54])) * x) + plom_powi(((pvd[54] * pvd[64]) * x), 2));
    }
    {
        const double x = obs[5];
        mean[5] = (pvd[55] * x);
        var[5] = (((pvd[55] * (1.0 - pvd[55])) * x) + plom_powi(((pvd[55] * pvd[65]) * x), 2));
    }
    {
        const double x = obs[6];
        mean[6] = (pvd[56] * x);
        var[6] = (((pvd[56] * (1.0 - pvd[56])) * x) + plom_powi(((pvd[56] * pvd[66]) * x), 2));
    }
    {
        const double x = obs[7];
        mean[7] = (pvd[57] * x);
        var[7] = (((pvd[57] * (1.0 - pvd[57])) * x) + plom_powi(((pvd[57] * pvd[67]) * x), 2));
    }
    {
        const double x = obs[8];
        mean[8] = (pvd[58] * x);
        var[8] = (((pvd[58] * (1.0 - pvd[58])) * x) + plom_powi(((pvd[58] * pvd[68]) * x), 2));
    }
    {
        const double x = obs[9];
        mean[9] = (pvd[59] * x);
        var[9] = (((pvd[59] * (1.0 - pvd[59])) * x) + plom_powi(((pvd[59] * pvd[69]) * x), 2));
    }
}

PLOM_HD void plom_seir10_hat_values(const double *X, const double *pvd, const double *cvd, const plom_drift_view *dm, double t, double *out)
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
    out[8] = X[8];
    out[9] = X[9];
    out[10] = X[10];
    out[11] = X[11];
    out[12] = X[12];
    out[13] = X[13];
    out[14] = X[14];
    out[15] = X[15];
    out[16] = X[16];
    out[17] = X[17];
    out[18] = X[18];
    out[19] = X[19];
    out[20] = X[20];
    out[21] = X[21];
    out[22] = X[22];
    out[23] = X[23];
    out[24] = X[24];
    out[25] = X[25];
    out[26] = X[26];
    out[27] = X[27];
    out[28] = X[28];
    out[29] = X[29];
    out[30] = cvd[0] - ((X[0] + X[10]) + X[20]);
    out[31] = cvd[1] - ((X[1] + X[11]) + X[21]);
    out[32] = cvd[2] - ((X[2] + X[12]) + X[22]);
    out[33] = cvd[3] - ((X[3] + X[13]) + X[23]);
    out[34] = cvd[4] - ((X[4] + X[14]) + X[24]);
    out[35] = cvd[5] - ((X[5] + X[15]) + X[25]);
    out[36] = cvd[6] - ((X[6] + X[16]) + X[26]);
    out[37] = cvd[7] - ((X[7] + X[17]) + X[27]);
    out[38] = cvd[8] - ((X[8] + X[18]) + X[28]);
    out[39] = cvd[9] - ((X[9] + X[19]) + X[29]);
    double obs[10], var[10];
    plom_seir10_proj2obs_f64(X, pvd, cvd, obs);
    plom_seir10_obs_moments(obs, pvd, cvd, t, out + 40, var);
}

PLOM_HD void plom_seir10_interval_psr_full_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
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
    const double P20 = PLOM_UNIFORM64(pv[20]);
    const double P21 = PLOM_UNIFORM64(pv[21]);
    const double P22 = PLOM_UNIFORM64(pv[22]);
    const double P23 = PLOM_UNIFORM64(pv[23]);
    const double P24 = PLOM_UNIFORM64(pv[24]);
    const double P25 = PLOM_UNIFORM64(pv[25]);
    const double P26 = PLOM_UNIFORM64(pv[26]);
    const double P27 = PLOM_UNIFORM64(pv[27]);
    const double P28 = PLOM_UNIFORM64(pv[28]);
    const double P29 = PLOM_UNIFORM64(pv[29]);
    const double P30 = PLOM_UNIFORM64(pv[30]);
    const double P31 = PLOM_UNIFORM64(pv[31]);
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
    const double P48 = PLOM_UNIFORM64(pv[48]);
    const double P49 = PLOM_UNIFORM64(pv[49]);
    const double P80 = PLOM_UNIFORM64(pv[80]);
    const double P81 = PLOM_UNIFORM64(pv[81]);
    const double P82 = PLOM_UNIFORM64(pv[82]);
    const double P83 = PLOM_UNIFORM64(pv[83]);
    const double P84 = PLOM_UNIFORM64(pv[84]);
    const double P85 = PLOM_UNIFORM64(pv[85]);
    const double P86 = PLOM_UNIFORM64(pv[86]);
    const double P87 = PLOM_UNIFORM64(pv[87]);
    const double P88 = PLOM_UNIFORM64(pv[88]);
    const double P89 = PLOM_UNIFORM64(pv[89]);
    const double P90 = PLOM_UNIFORM64(pv[90]);
    const double P91 = PLOM_UNIFORM64(pv[91]);
    const double P92 = PLOM_UNIFORM64(pv[92]);
    const double P93 = PLOM_UNIFORM64(pv[93]);
    const double P94 = PLOM_UNIFORM64(pv[94]);
    const double P95 = PLOM_UNIFORM64(pv[95]);
    const double P96 = PLOM_UNIFORM64(pv[96]);
    const double P97 = PLOM_UNIFORM64(pv[97]);
    const double P98 = PLOM_UNIFORM64(pv[98]);
    const double P99 = PLOM_UNIFORM64(pv[99]);
    const double P100 = PLOM_UNIFORM64(pv[100]);
    const double P101 = PLOM_UNIFORM64(pv[101]);
    const double P102 = PLOM_UNIFORM64(pv[102]);
    const double P103 = PLOM_UNIFORM64(pv[103]);
    const double P104 = PLOM_UNIFORM64(pv[104]);
    const double P105 = PLOM_UNIFORM64(pv[105]);
    const double P106 = PLOM_UNIFORM64(pv[106]);
    const double P107 = PLOM_UNIFORM64(pv[107]);
    const double P108 = PLOM_UNIFORM64(pv[108]);
    const double P109 = PLOM_UNIFORM64(pv[109]);
    const double P110 = PLOM_UNIFORM64(pv[110]);
    const double P111 = PLOM_UNIFORM64(pv[111]);
    const double P112 = PLOM_UNIFORM64(pv[112]);
    const double P113 = PLOM_UNIFORM64(pv[113]);
    const double P114 = PLOM_UNIFORM64(pv[114]);
    const double P115 = PLOM_UNIFORM64(pv[115]);
    const double P116 = PLOM_UNIFORM64(pv[116]);
    const double P117 = PLOM_UNIFORM64(pv[117]);
    const double P118 = PLOM_UNIFORM64(pv[118]);
    const double P119 = PLOM_UNIFORM64(pv[119]);
    const double P120 = PLOM_UNIFORM64(pv[120]);
    const double P121 = PLOM_UNIFORM64(pv[121]);
    const double P122 = PLOM_UNIFORM64(pv[122]);
    const double P123 = PLOM_UNIFORM64(pv[123]);
    const double P124 = PLOM_UNIFORM64(pv[124]);
    const double P125 = PLOM_UNIFORM64(pv[125]);
    const double P126 = PLOM_UNIFORM64(pv[126]);
    const double P127 = PLOM_UNIFORM64(pv[127]);
    const double P128 = PLOM_UNIFORM64(pv[128]);
    const double P129 = PLOM_UNIFORM64(pv[129]);
    const double P130 = PLOM_UNIFORM64(pv[130]);
    const double P131 = PLOM_UNIFORM64(pv[131]);
    const double P132 = PLOM_UNIFORM64(pv[132]);
    const double P133 = PLOM_UNIFORM64(pv[133]);
    const double P134 = PLOM_UNIFORM64(pv[134]);
    const double P135 = PLOM_UNIFORM64(pv[135]);
    const double P136 = PLOM_UNIFORM64(pv[136]);
    const double P137 = PLOM_UNIFORM64(pv[137]);
    const double P138 = PLOM_UNIFORM64(pv[138]);
    const double P139 = PLOM_UNIFORM64(pv[139]);
    const double P140 = PLOM_UNIFORM64(pv[140]);
    const double P141 = PLOM_UNIFORM64(pv[141]);
    const double P142 = PLOM_UNIFORM64(pv[142]);
    const double P143 = PLOM_UNIFORM64(pv[143]);
    const double P144 = PLOM_UNIFORM64(pv[144]);
    const double P145 = PLOM_UNIFORM64(pv[145]);
    const double P146 = PLOM_UNIFORM64(pv[146]);
    const double P147 = PLOM_UNIFORM64(pv[147]);
    const double P148 = PLOM_UNIFORM64(pv[148]);
    const double P149 = PLOM_UNIFORM64(pv[149]);
    const double P150 = PLOM_UNIFORM64(pv[150]);
    const double P151 = PLOM_UNIFORM64(pv[151]);
    const double P152 = PLOM_UNIFORM64(pv[152]);
    const double P153 = PLOM_UNIFORM64(pv[153]);
    const double P154 = PLOM_UNIFORM64(pv[154]);
    const double P155 = PLOM_UNIFORM64(pv[155]);
    const double P156 = PLOM_UNIFORM64(pv[156]);
    const double P157 = PLOM_UNIFORM64(pv[157]);
    const double P158 = PLOM_UNIFORM64(pv[158]);
    const double P159 = PLOM_UNIFORM64(pv[159]);
    const double P160 = PLOM_UNIFORM64(pv[160]);
    const double P161 = PLOM_UNIFORM64(pv[161]);
    const double P162 = PLOM_UNIFORM64(pv[162]);
    const double P163 = PLOM_UNIFORM64(pv[163]);
    const double P164 = PLOM_UNIFORM64(pv[164]);
    const double P165 = PLOM_UNIFORM64(pv[165]);
    const double P166 = PLOM_UNIFORM64(pv[166]);
    const double P167 = PLOM_UNIFORM64(pv[167]);
    const double P168 = PLOM_UNIFORM64(pv[168]);
    const double P169 = PLOM_UNIFORM64(pv[169]);
    const double P170 = PLOM_UNIFORM64(pv[170]);
    const double P171 = PLOM_UNIFORM64(pv[171]);
    const double P172 = PLOM_UNIFORM64(pv[172]);
    const double P173 = PLOM_UNIFORM64(pv[173]);
    const double P174 = PLOM_UNIFORM64(pv[174]);
    const double P175 = PLOM_UNIFORM64(pv[175]);
    const double P176 = PLOM_UNIFORM64(pv[176]);
    const double P177 = PLOM_UNIFORM64(pv[177]);
    const double P178 = PLOM_UNIFORM64(pv[178]);
    const double P179 = PLOM_UNIFORM64(pv[179]);
    X[30] = 0.0; /* reset_inc */
    X[31] = 0.0; /* reset_inc */
    X[32] = 0.0; /* reset_inc */
    X[33] = 0.0; /* reset_inc */
    X[34] = 0.0; /* reset_inc */
    X[35] = 0.0; /* reset_inc */
    X[36] = 0.0; /* reset_inc */
    X[37] = 0.0; /* reset_inc */
    X[38] = 0.0; /* reset_inc */
    X[39] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double C4 = plom_cov_ld(cvk + 4);
        const double C5 = plom_cov_ld(cvk + 5);
        const double C6 = plom_cov_ld(cvk + 6);
        const double C7 = plom_cov_ld(cvk + 7);
        const double C8 = plom_cov_ld(cvk + 8);
        const double C9 = plom_cov_ld(cvk + 9);
        double n0_0 = 0.0;
        double n1_0 = 0.0;
        double n2_0 = 0.0;
        double n3_0 = 0.0;
        double n0_1 = 0.0;
        double n1_1 = 0.0;
        double n2_1 = 0.0;
        double n3_1 = 0.0;
        double n0_2 = 0.0;
        double n1_2 = 0.0;
        double n2_2 = 0.0;
        double n3_2 = 0.0;
        double n0_3 = 0.0;
        double n1_3 = 0.0;
        double n2_3 = 0.0;
        double n3_3 = 0.0;
        double n0_4 = 0.0;
        double n1_4 = 0.0;
        double n2_4 = 0.0;
        double n3_4 = 0.0;
        double n0_5 = 0.0;
        double n1_5 = 0.0;
        double n2_5 = 0.0;
        double n3_5 = 0.0;
        double n0_6 = 0.0;
        double n1_6 = 0.0;
        double n2_6 = 0.0;
        double n3_6 = 0.0;
        double n0_7 = 0.0;
        double n1_7 = 0.0;
        double n2_7 = 0.0;
        double n3_7 = 0.0;
        double n0_8 = 0.0;
        double n1_8 = 0.0;
        double n2_8 = 0.0;
        double n3_8 = 0.0;
        double n0_9 = 0.0;
        double n1_9 = 0.0;
        double n2_9 = 0.0;
        double n3_9 = 0.0;
        double stay0 = 0.0;
        double stay1 = 0.0;
        double stay2 = 0.0;
        double stay3 = 0.0;
        double stay4 = 0.0;
        double stay5 = 0.0;
        double stay6 = 0.0;
        double stay7 = 0.0;
        double stay8 = 0.0;
        double stay9 = 0.0;
        double stay10 = 0.0;
        double stay11 = 0.0;
        double stay12 = 0.0;
        double stay13 = 0.0;
        double stay14 = 0.0;
        double stay15 = 0.0;
        double stay16 = 0.0;
        double stay17 = 0.0;
        double stay18 = 0.0;
        double stay19 = 0.0;
        double stay20 = 0.0;
        double stay21 = 0.0;
        double stay22 = 0.0;
        double stay23 = 0.0;
        double stay24 = 0.0;
        double stay25 = 0.0;
        double stay26 = 0.0;
        double stay27 = 0.0;
        double stay28 = 0.0;
        double stay29 = 0.0;
        (void)n0_0; (void)n1_0; (void)n2_0; (void)n3_0; (void)n0_1; (void)n1_1; (void)n2_1; (void)n3_1; (void)n0_2; (void)n1_2; (void)n2_2; (void)n3_2; (void)n0_3; (void)n1_3; (void)n2_3; (void)n3_3; (void)n0_4; (void)n1_4; (void)n2_4; (void)n3_4; (void)n0_5; (void)n1_5; (void)n2_5; (void)n3_5; (void)n0_6; (void)n1_6; (void)n2_6; (void)n3_6; (void)n0_7; (void)n1_7; (void)n2_7; (void)n3_7; (void)n0_8; (void)n1_8; (void)n2_8; (void)n3_8; (void)n0_9; (void)n1_9; (void)n2_9; (void)n3_9;
        const double r0_0 = ((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])) / C0);
        const double r1_0 = P10;
        const double r2_0 = P20;
        const double r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        const double r0_1 = ((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])) / C1);
        const double r1_1 = P11;
        const double r2_1 = P21;
        const double r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        const double r0_2 = ((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])) / C2);
        const double r1_2 = P12;
        const double r2_2 = P22;
        const double r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        const double r0_3 = ((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])) / C3);
        const double r1_3 = P13;
        const double r2_3 = P23;
        const double r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        const double r0_4 = ((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])) / C4);
        const double r1_4 = P14;
        const double r2_4 = P24;
        const double r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        const double r0_5 = ((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])) / C5);
        const double r1_5 = P15;
        const double r2_5 = P25;
        const double r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        const double r0_6 = ((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])) / C6);
        const double r1_6 = P16;
        const double r2_6 = P26;
        const double r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        const double r0_7 = ((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])) / C7);
        const double r1_7 = P17;
        const double r2_7 = P27;
        const double r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        const double r0_8 = ((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])) / C8);
        const double r1_8 = P18;
        const double r2_8 = P28;
        const double r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        const double r0_9 = ((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])) / C9);
        const double r1_9 = P19;
        const double r2_9 = P29;
        const double r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        const double sd0_0 = P40;
        const double v0_0 = sd0_0 * sd0_0;
        const double xi0_0 = (v0_0 * plom_gamma_draw(ns, dt / v0_0)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_0 * xi0_0) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_0 * xi0_0) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[0]), p, cnt);
            n0_0 = cnt[0];
            stay0 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_0 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_0 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[10]), p, cnt);
            n1_0 = cnt[0];
            stay10 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_0 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_0 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[20]), p, cnt);
            n2_0 = cnt[0];
            stay20 = cnt[1];
        }
        n3_0 = plom_poisson_draw(ns, r3_0 * dt);
        const double sd0_1 = P41;
        const double v0_1 = sd0_1 * sd0_1;
        const double xi0_1 = (v0_1 * plom_gamma_draw(ns, dt / v0_1)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_1 * xi0_1) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_1 * xi0_1) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[1]), p, cnt);
            n0_1 = cnt[0];
            stay1 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_1 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_1 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[11]), p, cnt);
            n1_1 = cnt[0];
            stay11 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_1 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_1 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[21]), p, cnt);
            n2_1 = cnt[0];
            stay21 = cnt[1];
        }
        n3_1 = plom_poisson_draw(ns, r3_1 * dt);
        const double sd0_2 = P42;
        const double v0_2 = sd0_2 * sd0_2;
        const double xi0_2 = (v0_2 * plom_gamma_draw(ns, dt / v0_2)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_2 * xi0_2) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_2 * xi0_2) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[2]), p, cnt);
            n0_2 = cnt[0];
            stay2 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_2 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_2 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[12]), p, cnt);
            n1_2 = cnt[0];
            stay12 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_2 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_2 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[22]), p, cnt);
            n2_2 = cnt[0];
            stay22 = cnt[1];
        }
        n3_2 = plom_poisson_draw(ns, r3_2 * dt);
        const double sd0_3 = P43;
        const double v0_3 = sd0_3 * sd0_3;
        const double xi0_3 = (v0_3 * plom_gamma_draw(ns, dt / v0_3)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_3 * xi0_3) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_3 * xi0_3) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[3]), p, cnt);
            n0_3 = cnt[0];
            stay3 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_3 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_3 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[13]), p, cnt);
            n1_3 = cnt[0];
            stay13 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_3 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_3 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[23]), p, cnt);
            n2_3 = cnt[0];
            stay23 = cnt[1];
        }
        n3_3 = plom_poisson_draw(ns, r3_3 * dt);
        const double sd0_4 = P44;
        const double v0_4 = sd0_4 * sd0_4;
        const double xi0_4 = (v0_4 * plom_gamma_draw(ns, dt / v0_4)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_4 * xi0_4) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_4 * xi0_4) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[4]), p, cnt);
            n0_4 = cnt[0];
            stay4 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_4 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_4 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[14]), p, cnt);
            n1_4 = cnt[0];
            stay14 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_4 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_4 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[24]), p, cnt);
            n2_4 = cnt[0];
            stay24 = cnt[1];
        }
        n3_4 = plom_poisson_draw(ns, r3_4 * dt);
        const double sd0_5 = P45;
        const double v0_5 = sd0_5 * sd0_5;
        const double xi0_5 = (v0_5 * plom_gamma_draw(ns, dt / v0_5)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_5 * xi0_5) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_5 * xi0_5) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[5]), p, cnt);
            n0_5 = cnt[0];
            stay5 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_5 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_5 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[15]), p, cnt);
            n1_5 = cnt[0];
            stay15 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_5 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_5 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[25]), p, cnt);
            n2_5 = cnt[0];
            stay25 = cnt[1];
        }
        n3_5 = plom_poisson_draw(ns, r3_5 * dt);
        const double sd0_6 = P46;
        const double v0_6 = sd0_6 * sd0_6;
        const double xi0_6 = (v0_6 * plom_gamma_draw(ns, dt / v0_6)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_6 * xi0_6) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_6 * xi0_6) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[6]), p, cnt);
            n0_6 = cnt[0];
            stay6 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_6 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_6 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[16]), p, cnt);
            n1_6 = cnt[0];
            stay16 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_6 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_6 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[26]), p, cnt);
            n2_6 = cnt[0];
            stay26 = cnt[1];
        }
        n3_6 = plom_poisson_draw(ns, r3_6 * dt);
        const double sd0_7 = P47;
        const double v0_7 = sd0_7 * sd0_7;
        const double xi0_7 = (v0_7 * plom_gamma_draw(ns, dt / v0_7)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_7 * xi0_7) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_7 * xi0_7) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[7]), p, cnt);
            n0_7 = cnt[0];
            stay7 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_7 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_7 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[17]), p, cnt);
            n1_7 = cnt[0];
            stay17 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_7 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_7 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[27]), p, cnt);
            n2_7 = cnt[0];
            stay27 = cnt[1];
        }
        n3_7 = plom_poisson_draw(ns, r3_7 * dt);
        const double sd0_8 = P48;
        const double v0_8 = sd0_8 * sd0_8;
        const double xi0_8 = (v0_8 * plom_gamma_draw(ns, dt / v0_8)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_8 * xi0_8) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_8 * xi0_8) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[8]), p, cnt);
            n0_8 = cnt[0];
            stay8 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_8 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_8 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[18]), p, cnt);
            n1_8 = cnt[0];
            stay18 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_8 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_8 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[28]), p, cnt);
            n2_8 = cnt[0];
            stay28 = cnt[1];
        }
        n3_8 = plom_poisson_draw(ns, r3_8 * dt);
        const double sd0_9 = P49;
        const double v0_9 = sd0_9 * sd0_9;
        const double xi0_9 = (v0_9 * plom_gamma_draw(ns, dt / v0_9)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_9 * xi0_9) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_9 * xi0_9) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[9]), p, cnt);
            n0_9 = cnt[0];
            stay9 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_9 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_9 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[19]), p, cnt);
            n1_9 = cnt[0];
            stay19 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_9 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_9 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[29]), p, cnt);
            n2_9 = cnt[0];
            stay29 = cnt[1];
        }
        n3_9 = plom_poisson_draw(ns, r3_9 * dt);
        X[0] = stay0 + n3_0;
        X[10] = stay10 + n0_0;
        X[20] = stay20 + n1_0;
        X[1] = stay1 + n3_1;
        X[11] = stay11 + n0_1;
        X[21] = stay21 + n1_1;
        X[2] = stay2 + n3_2;
        X[12] = stay12 + n0_2;
        X[22] = stay22 + n1_2;
        X[3] = stay3 + n3_3;
        X[13] = stay13 + n0_3;
        X[23] = stay23 + n1_3;
        X[4] = stay4 + n3_4;
        X[14] = stay14 + n0_4;
        X[24] = stay24 + n1_4;
        X[5] = stay5 + n3_5;
        X[15] = stay15 + n0_5;
        X[25] = stay25 + n1_5;
        X[6] = stay6 + n3_6;
        X[16] = stay16 + n0_6;
        X[26] = stay26 + n1_6;
        X[7] = stay7 + n3_7;
        X[17] = stay17 + n0_7;
        X[27] = stay27 + n1_7;
        X[8] = stay8 + n3_8;
        X[18] = stay18 + n0_8;
        X[28] = stay28 + n1_8;
        X[9] = stay9 + n3_9;
        X[19] = stay19 + n0_9;
        X[29] = stay29 + n1_9;
        X[30] = X[30] + (n1_0);
        X[31] = X[31] + (n1_1);
        X[32] = X[32] + (n1_2);
        X[33] = X[33] + (n1_3);
        X[34] = X[34] + (n1_4);
        X[35] = X[35] + (n1_5);
        X[36] = X[36] + (n1_6);
        X[37] = X[37] + (n1_7);
        X[38] = X[38] + (n1_8);
        X[39] = X[39] + (n1_9);
        t += dt;
    }
}

PLOM_HD void plom_seir10_interval_psr_no_env_sto_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
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
    const double P20 = PLOM_UNIFORM64(pv[20]);
    const double P21 = PLOM_UNIFORM64(pv[21]);
    const double P22 = PLOM_UNIFORM64(pv[22]);
    const double P23 = PLOM_UNIFORM64(pv[23]);
    const double P24 = PLOM_UNIFORM64(pv[24]);
    const double P25 = PLOM_UNIFORM64(pv[25]);
    const double P26 = PLOM_UNIFORM64(pv[26]);
    const double P27 = PLOM_UNIFORM64(pv[27]);
    const double P28 = PLOM_UNIFORM64(pv[28]);
    const double P29 = PLOM_UNIFORM64(pv[29]);
    const double P30 = PLOM_UNIFORM64(pv[30]);
    const double P31 = PLOM_UNIFORM64(pv[31]);
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P80 = PLOM_UNIFORM64(pv[80]);
    const double P81 = PLOM_UNIFORM64(pv[81]);
    const double P82 = PLOM_UNIFORM64(pv[82]);
    const double P83 = PLOM_UNIFORM64(pv[83]);
    const double P84 = PLOM_UNIFORM64(pv[84]);
    const double P85 = PLOM_UNIFORM64(pv[85]);
    const double P86 = PLOM_UNIFORM64(pv[86]);
    const double P87 = PLOM_UNIFORM64(pv[87]);
    const double P88 = PLOM_UNIFORM64(pv[88]);
    const double P89 = PLOM_UNIFORM64(pv[89]);
    const double P90 = PLOM_UNIFORM64(pv[90]);
    const double P91 = PLOM_UNIFORM64(pv[91]);
    const double P92 = PLOM_UNIFORM64(pv[92]);
    const double P93 = PLOM_UNIFORM64(pv[93]);
    const double P94 = PLOM_UNIFORM64(pv[94]);
    const double P95 = PLOM_UNIFORM64(pv[95]);
    const double P96 = PLOM_UNIFORM64(pv[96]);
    const double P97 = PLOM_UNIFORM64(pv[97]);
    const double P98 = PLOM_UNIFORM64(pv[98]);
    const double P99 = PLOM_UNIFORM64(pv[99]);
    const double P100 = PLOM_UNIFORM64(pv[100]);
    const double P101 = PLOM_UNIFORM64(pv[101]);
    const double P102 = PLOM_UNIFORM64(pv[102]);
    const double P103 = PLOM_UNIFORM64(pv[103]);
    const double P104 = PLOM_UNIFORM64(pv[104]);
    const double P105 = PLOM_UNIFORM64(pv[105]);
    const double P106 = PLOM_UNIFORM64(pv[106]);
    const double P107 = PLOM_UNIFORM64(pv[107]);
    const double P108 = PLOM_UNIFORM64(pv[108]);
    const double P109 = PLOM_UNIFORM64(pv[109]);
    const double P110 = PLOM_UNIFORM64(pv[110]);
    const double P111 = PLOM_UNIFORM64(pv[111]);
    const double P112 = PLOM_UNIFORM64(pv[112]);
    const double P113 = PLOM_UNIFORM64(pv[113]);
    const double P114 = PLOM_UNIFORM64(pv[114]);
    const double P115 = PLOM_UNIFORM64(pv[115]);
    const double P116 = PLOM_UNIFORM64(pv[116]);
    const double P117 = PLOM_UNIFORM64(pv[117]);
    const double P118 = PLOM_UNIFORM64(pv[118]);
    const double P119 = PLOM_UNIFORM64(pv[119]);
    const double P120 = PLOM_UNIFORM64(pv[120]);
    const double P121 = PLOM_UNIFORM64(pv[121]);
    const double P122 = PLOM_UNIFORM64(pv[122]);
    const double P123 = PLOM_UNIFORM64(pv[123]);
    const double P124 = PLOM_UNIFORM64(pv[124]);
    const double P125 = PLOM_UNIFORM64(pv[125]);
    const double P126 = PLOM_UNIFORM64(pv[126]);
    const double P127 = PLOM_UNIFORM64(pv[127]);
    const double P128 = PLOM_UNIFORM64(pv[128]);
    const double P129 = PLOM_UNIFORM64(pv[129]);
    const double P130 = PLOM_UNIFORM64(pv[130]);
    const double P131 = PLOM_UNIFORM64(pv[131]);
    const double P132 = PLOM_UNIFORM64(pv[132]);
    const double P133 = PLOM_UNIFORM64(pv[133]);
    const double P134 = PLOM_UNIFORM64(pv[134]);
    const double P135 = PLOM_UNIFORM64(pv[135]);
    const double P136 = PLOM_UNIFORM64(pv[136]);
    const double P137 = PLOM_UNIFORM64(pv[137]);
    const double P138 = PLOM_UNIFORM64(pv[138]);
    const double P139 = PLOM_UNIFORM64(pv[139]);
    const double P140 = PLOM_UNIFORM64(pv[140]);
    const double P141 = PLOM_UNIFORM64(pv[141]);
    const double P142 = PLOM_UNIFORM64(pv[142]);
    const double P143 = PLOM_UNIFORM64(pv[143]);
    const double P144 = PLOM_UNIFORM64(pv[144]);
    const double P145 = PLOM_UNIFORM64(pv[145]);
    const double P146 = PLOM_UNIFORM64(pv[146]);
    const double P147 = PLOM_UNIFORM64(pv[147]);
    const double P148 = PLOM_UNIFORM64(pv[148]);
    const double P149 = PLOM_UNIFORM64(pv[149]);
    const double P150 = PLOM_UNIFORM64(pv[150]);
    const double P151 = PLOM_UNIFORM64(pv[151]);
    const double P152 = PLOM_UNIFORM64(pv[152]);
    const double P153 = PLOM_UNIFORM64(pv[153]);
    const double P154 = PLOM_UNIFORM64(pv[154]);
    const double P155 = PLOM_UNIFORM64(pv[155]);
    const double P156 = PLOM_UNIFORM64(pv[156]);
    const double P157 = PLOM_UNIFORM64(pv[157]);
    const double P158 = PLOM_UNIFORM64(pv[158]);
    const double P159 = PLOM_UNIFORM64(pv[159]);
    const double P160 = PLOM_UNIFORM64(pv[160]);
    const double P161 = PLOM_UNIFORM64(pv[161]);
    const double P162 = PLOM_UNIFORM64(pv[162]);
    const double P163 = PLOM_UNIFORM64(pv[163]);
    const double P164 = PLOM_UNIFORM64(pv[164]);
    const double P165 = PLOM_UNIFORM64(pv[165]);
    const double P166 = PLOM_UNIFORM64(pv[166]);
    const double P167 = PLOM_UNIFORM64(pv[167]);
    const double P168 = PLOM_UNIFORM64(pv[168]);
    const double P169 = PLOM_UNIFORM64(pv[169]);
    const double P170 = PLOM_UNIFORM64(pv[170]);
    const double P171 = PLOM_UNIFORM64(pv[171]);
    const double P172 = PLOM_UNIFORM64(pv[172]);
    const double P173 = PLOM_UNIFORM64(pv[173]);
    const double P174 = PLOM_UNIFORM64(pv[174]);
    const double P175 = PLOM_UNIFORM64(pv[175]);
    const double P176 = PLOM_UNIFORM64(pv[176]);
    const double P177 = PLOM_UNIFORM64(pv[177]);
    const double P178 = PLOM_UNIFORM64(pv[178]);
    const double P179 = PLOM_UNIFORM64(pv[179]);
    X[30] = 0.0; /* reset_inc */
    X[31] = 0.0; /* reset_inc */
    X[32] = 0.0; /* reset_inc */
    X[33] = 0.0; /* reset_inc */
    X[34] = 0.0; /* reset_inc */
    X[35] = 0.0; /* reset_inc */
    X[36] = 0.0; /* reset_inc */
    X[37] = 0.0; /* reset_inc */
    X[38] = 0.0; /* reset_inc */
    X[39] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double C4 = plom_cov_ld(cvk + 4);
        const double C5 = plom_cov_ld(cvk + 5);
        const double C6 = plom_cov_ld(cvk + 6);
        const double C7 = plom_cov_ld(cvk + 7);
        const double C8 = plom_cov_ld(cvk + 8);
        const double C9 = plom_cov_ld(cvk + 9);
        double n0_0 = 0.0;
        double n1_0 = 0.0;
        double n2_0 = 0.0;
        double n3_0 = 0.0;
        double n0_1 = 0.0;
        double n1_1 = 0.0;
        double n2_1 = 0.0;
        double n3_1 = 0.0;
        double n0_2 = 0.0;
        double n1_2 = 0.0;
        double n2_2 = 0.0;
        double n3_2 = 0.0;
        double n0_3 = 0.0;
        double n1_3 = 0.0;
        double n2_3 = 0.0;
        double n3_3 = 0.0;
        double n0_4 = 0.0;
        double n1_4 = 0.0;
        double n2_4 = 0.0;
        double n3_4 = 0.0;
        double n0_5 = 0.0;
        double n1_5 = 0.0;
        double n2_5 = 0.0;
        double n3_5 = 0.0;
        double n0_6 = 0.0;
        double n1_6 = 0.0;
        double n2_6 = 0.0;
        double n3_6 = 0.0;
        double n0_7 = 0.0;
        double n1_7 = 0.0;
        double n2_7 = 0.0;
        double n3_7 = 0.0;
        double n0_8 = 0.0;
        double n1_8 = 0.0;
        double n2_8 = 0.0;
        double n3_8 = 0.0;
        double n0_9 = 0.0;
        double n1_9 = 0.0;
        double n2_9 = 0.0;
        double n3_9 = 0.0;
        double stay0 = 0.0;
        double stay1 = 0.0;
        double stay2 = 0.0;
        double stay3 = 0.0;
        double stay4 = 0.0;
        double stay5 = 0.0;
        double stay6 = 0.0;
        double stay7 = 0.0;
        double stay8 = 0.0;
        double stay9 = 0.0;
        double stay10 = 0.0;
        double stay11 = 0.0;
        double stay12 = 0.0;
        double stay13 = 0.0;
        double stay14 = 0.0;
        double stay15 = 0.0;
        double stay16 = 0.0;
        double stay17 = 0.0;
        double stay18 = 0.0;
        double stay19 = 0.0;
        double stay20 = 0.0;
        double stay21 = 0.0;
        double stay22 = 0.0;
        double stay23 = 0.0;
        double stay24 = 0.0;
        double stay25 = 0.0;
        double stay26 = 0.0;
        double stay27 = 0.0;
        double stay28 = 0.0;
        double stay29 = 0.0;
        (void)n0_0; (void)n1_0; (void)n2_0; (void)n3_0; (void)n0_1; (void)n1_1; (void)n2_1; (void)n3_1; (void)n0_2; (void)n1_2; (void)n2_2; (void)n3_2; (void)n0_3; (void)n1_3; (void)n2_3; (void)n3_3; (void)n0_4; (void)n1_4; (void)n2_4; (void)n3_4; (void)n0_5; (void)n1_5; (void)n2_5; (void)n3_5; (void)n0_6; (void)n1_6; (void)n2_6; (void)n3_6; (void)n0_7; (void)n1_7; (void)n2_7; (void)n3_7; (void)n0_8; (void)n1_8; (void)n2_8; (void)n3_8; (void)n0_9; (void)n1_9; (void)n2_9; (void)n3_9;
        const double r0_0 = ((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])) / C0);
        const double r1_0 = P10;
        const double r2_0 = P20;
        const double r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        const double r0_1 = ((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])) / C1);
        const double r1_1 = P11;
        const double r2_1 = P21;
        const double r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        const double r0_2 = ((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])) / C2);
        const double r1_2 = P12;
        const double r2_2 = P22;
        const double r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        const double r0_3 = ((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])) / C3);
        const double r1_3 = P13;
        const double r2_3 = P23;
        const double r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        const double r0_4 = ((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])) / C4);
        const double r1_4 = P14;
        const double r2_4 = P24;
        const double r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        const double r0_5 = ((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])) / C5);
        const double r1_5 = P15;
        const double r2_5 = P25;
        const double r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        const double r0_6 = ((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])) / C6);
        const double r1_6 = P16;
        const double r2_6 = P26;
        const double r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        const double r0_7 = ((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])) / C7);
        const double r1_7 = P17;
        const double r2_7 = P27;
        const double r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        const double r0_8 = ((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])) / C8);
        const double r1_8 = P18;
        const double r2_8 = P28;
        const double r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        const double r0_9 = ((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])) / C9);
        const double r1_9 = P19;
        const double r2_9 = P29;
        const double r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        const double xi0_0 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_0 * xi0_0) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_0 * xi0_0) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[0]), p, cnt);
            n0_0 = cnt[0];
            stay0 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_0 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_0 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[10]), p, cnt);
            n1_0 = cnt[0];
            stay10 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_0 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_0 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[20]), p, cnt);
            n2_0 = cnt[0];
            stay20 = cnt[1];
        }
        n3_0 = plom_poisson_draw(ns, r3_0 * dt);
        const double xi0_1 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_1 * xi0_1) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_1 * xi0_1) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[1]), p, cnt);
            n0_1 = cnt[0];
            stay1 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_1 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_1 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[11]), p, cnt);
            n1_1 = cnt[0];
            stay11 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_1 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_1 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[21]), p, cnt);
            n2_1 = cnt[0];
            stay21 = cnt[1];
        }
        n3_1 = plom_poisson_draw(ns, r3_1 * dt);
        const double xi0_2 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_2 * xi0_2) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_2 * xi0_2) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[2]), p, cnt);
            n0_2 = cnt[0];
            stay2 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_2 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_2 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[12]), p, cnt);
            n1_2 = cnt[0];
            stay12 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_2 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_2 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[22]), p, cnt);
            n2_2 = cnt[0];
            stay22 = cnt[1];
        }
        n3_2 = plom_poisson_draw(ns, r3_2 * dt);
        const double xi0_3 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_3 * xi0_3) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_3 * xi0_3) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[3]), p, cnt);
            n0_3 = cnt[0];
            stay3 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_3 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_3 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[13]), p, cnt);
            n1_3 = cnt[0];
            stay13 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_3 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_3 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[23]), p, cnt);
            n2_3 = cnt[0];
            stay23 = cnt[1];
        }
        n3_3 = plom_poisson_draw(ns, r3_3 * dt);
        const double xi0_4 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_4 * xi0_4) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_4 * xi0_4) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[4]), p, cnt);
            n0_4 = cnt[0];
            stay4 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_4 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_4 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[14]), p, cnt);
            n1_4 = cnt[0];
            stay14 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_4 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_4 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[24]), p, cnt);
            n2_4 = cnt[0];
            stay24 = cnt[1];
        }
        n3_4 = plom_poisson_draw(ns, r3_4 * dt);
        const double xi0_5 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_5 * xi0_5) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_5 * xi0_5) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[5]), p, cnt);
            n0_5 = cnt[0];
            stay5 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_5 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_5 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[15]), p, cnt);
            n1_5 = cnt[0];
            stay15 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_5 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_5 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[25]), p, cnt);
            n2_5 = cnt[0];
            stay25 = cnt[1];
        }
        n3_5 = plom_poisson_draw(ns, r3_5 * dt);
        const double xi0_6 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_6 * xi0_6) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_6 * xi0_6) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[6]), p, cnt);
            n0_6 = cnt[0];
            stay6 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_6 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_6 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[16]), p, cnt);
            n1_6 = cnt[0];
            stay16 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_6 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_6 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[26]), p, cnt);
            n2_6 = cnt[0];
            stay26 = cnt[1];
        }
        n3_6 = plom_poisson_draw(ns, r3_6 * dt);
        const double xi0_7 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_7 * xi0_7) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_7 * xi0_7) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[7]), p, cnt);
            n0_7 = cnt[0];
            stay7 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_7 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_7 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[17]), p, cnt);
            n1_7 = cnt[0];
            stay17 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_7 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_7 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[27]), p, cnt);
            n2_7 = cnt[0];
            stay27 = cnt[1];
        }
        n3_7 = plom_poisson_draw(ns, r3_7 * dt);
        const double xi0_8 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_8 * xi0_8) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_8 * xi0_8) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[8]), p, cnt);
            n0_8 = cnt[0];
            stay8 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_8 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_8 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[18]), p, cnt);
            n1_8 = cnt[0];
            stay18 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_8 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_8 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[28]), p, cnt);
            n2_8 = cnt[0];
            stay28 = cnt[1];
        }
        n3_8 = plom_poisson_draw(ns, r3_8 * dt);
        const double xi0_9 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_9 * xi0_9) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_9 * xi0_9) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[9]), p, cnt);
            n0_9 = cnt[0];
            stay9 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_9 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_9 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[19]), p, cnt);
            n1_9 = cnt[0];
            stay19 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_9 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_9 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[29]), p, cnt);
            n2_9 = cnt[0];
            stay29 = cnt[1];
        }
        n3_9 = plom_poisson_draw(ns, r3_9 * dt);
        X[0] = stay0 + n3_0;
        X[10] = stay10 + n0_0;
        X[20] = stay20 + n1_0;
        X[1] = stay1 + n3_1;
        X[11] = stay11 + n0_1;
        X[21] = stay21 + n1_1;
        X[2] = stay2 + n3_2;
        X[12] = stay12 + n0_2;
        X[22] = stay22 + n1_2;
        X[3] = stay3 + n3_3;
        X[13] = stay13 + n0_3;
        X[23] = stay23 + n1_3;
        X[4] = stay4 + n3_4;
        X[14] = stay14 + n0_4;
        X[24] = stay24 + n1_4;
        X[5] = stay5 + n3_5;
        X[15] = stay15 + n0_5;
        X[25] = stay25 + n1_5;
        X[6] = stay6 + n3_6;
        X[16] = stay16 + n0_6;
        X[26] = stay26 + n1_6;
        X[7] = stay7 + n3_7;
        X[17] = stay17 + n0_7;
        X[27] = stay27 + n1_7;
        X[8] = stay8 + n3_8;
        X[18] = stay18 + n0_8;
        X[28] = stay28 + n1_8;
        X[9] = stay9 + n3_9;
        X[19] = stay19 + n0_9;
        X[29] = stay29 + n1_9;
        X[30] = X[30] + (n1_0);
        X[31] = X[31] + (n1_1);
        X[32] = X[32] + (n1_2);
        X[33] = X[33] + (n1_3);
        X[34] = X[34] + (n1_4);
        X[35] = X[35] + (n1_5);
        X[36] = X[36] + (n1_6);
        X[37] = X[37] + (n1_7);
        X[38] = X[38] + (n1_8);
        X[39] = X[39] + (n1_9);
        t += dt;
    }
}

PLOM_HD void plom_seir10_ode_rhs(double t, const double *X, const double *pv, const double *lp, const double *cv, const plom_spline_set *sp, double *f)
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
    const double P16 = PLOM_UNIFORM64(pv[16]);
    const double P17 = PLOM_UNIFORM64(pv[17]);
    const double P18 = PLOM_UNIFORM64(pv[18]);
    const double P19 = PLOM_UNIFORM64(pv[19]);
    const double P20 = PLOM_UNIFORM64(pv[20]);
    const double P21 = PLOM_UNIFORM64(pv[21]);
    const double P22 = PLOM_UNIFORM64(pv[22]);
    const double P23 = PLOM_UNIFORM64(pv[23]);
    const double P24 = PLOM_UNIFORM64(pv[24]);
    const double P25 = PLOM_UNIFORM64(pv[25]);
    const double P26 = PLOM_UNIFORM64(pv[26]);
    const double P27 = PLOM_UNIFORM64(pv[27]);
    const double P28 = PLOM_UNIFORM64(pv[28]);
    const double P29 = PLOM_UNIFORM64(pv[29]);
    const double P30 = PLOM_UNIFORM64(pv[30]);
    const double P31 = PLOM_UNIFORM64(pv[31]);
    const double P32 = PLOM_UNIFORM64(pv[32]);
    const double P33 = PLOM_UNIFORM64(pv[33]);
    const double P34 = PLOM_UNIFORM64(pv[34]);
    const double P35 = PLOM_UNIFORM64(pv[35]);
    const double P36 = PLOM_UNIFORM64(pv[36]);
    const double P37 = PLOM_UNIFORM64(pv[37]);
    const double P38 = PLOM_UNIFORM64(pv[38]);
    const double P39 = PLOM_UNIFORM64(pv[39]);
    const double P80 = PLOM_UNIFORM64(pv[80]);
    const double P81 = PLOM_UNIFORM64(pv[81]);
    const double P82 = PLOM_UNIFORM64(pv[82]);
    const double P83 = PLOM_UNIFORM64(pv[83]);
    const double P84 = PLOM_UNIFORM64(pv[84]);
    const double P85 = PLOM_UNIFORM64(pv[85]);
    const double P86 = PLOM_UNIFORM64(pv[86]);
    const double P87 = PLOM_UNIFORM64(pv[87]);
    const double P88 = PLOM_UNIFORM64(pv[88]);
    const double P89 = PLOM_UNIFORM64(pv[89]);
    const double P90 = PLOM_UNIFORM64(pv[90]);
    const double P91 = PLOM_UNIFORM64(pv[91]);
    const double P92 = PLOM_UNIFORM64(pv[92]);
    const double P93 = PLOM_UNIFORM64(pv[93]);
    const double P94 = PLOM_UNIFORM64(pv[94]);
    const double P95 = PLOM_UNIFORM64(pv[95]);
    const double P96 = PLOM_UNIFORM64(pv[96]);
    const double P97 = PLOM_UNIFORM64(pv[97]);
    const double P98 = PLOM_UNIFORM64(pv[98]);
    const double P99 = PLOM_UNIFORM64(pv[99]);
    const double P100 = PLOM_UNIFORM64(pv[100]);
    const double P101 = PLOM_UNIFORM64(pv[101]);
    const double P102 = PLOM_UNIFORM64(pv[102]);
    const double P103 = PLOM_UNIFORM64(pv[103]);
    const double P104 = PLOM_UNIFORM64(pv[104]);
    const double P105 = PLOM_UNIFORM64(pv[105]);
    const double P106 = PLOM_UNIFORM64(pv[106]);
    const double P107 = PLOM_UNIFORM64(pv[107]);
    const double P108 = PLOM_UNIFORM64(pv[108]);
    const double P109 = PLOM_UNIFORM64(pv[109]);
    const double P110 = PLOM_UNIFORM64(pv[110]);
    const double P111 = PLOM_UNIFORM64(pv[111]);
    const double P112 = PLOM_UNIFORM64(pv[112]);
    const double P113 = PLOM_UNIFORM64(pv[113]);
    const double P114 = PLOM_UNIFORM64(pv[114]);
    const double P115 = PLOM_UNIFORM64(pv[115]);
    const double P116 = PLOM_UNIFORM64(pv[116]);
    const double P117 = PLOM_UNIFORM64(pv[117]);
    const double P118 = PLOM_UNIFORM64(pv[118]);
    const double P119 = PLOM_UNIFORM64(pv[119]);
    const double P120 = PLOM_UNIFORM64(pv[120]);
    const double P121 = PLOM_UNIFORM64(pv[121]);
    const double P122 = PLOM_UNIFORM64(pv[122]);
    const double P123 = PLOM_UNIFORM64(pv[123]);
    const double P124 = PLOM_UNIFORM64(pv[124]);
    const double P125 = PLOM_UNIFORM64(pv[125]);
    const double P126 = PLOM_UNIFORM64(pv[126]);
    const double P127 = PLOM_UNIFORM64(pv[127]);
    const double P128 = PLOM_UNIFORM64(pv[128]);
    const double P129 = PLOM_UNIFORM64(pv[129]);
    const double P130 = PLOM_UNIFORM64(pv[130]);
    const double P131 = PLOM_UNIFORM64(pv[131]);
    const double P132 = PLOM_UNIFORM64(pv[132]);
    const double P133 = PLOM_UNIFORM64(pv[133]);
    const double P134 = PLOM_UNIFORM64(pv[134]);
    const double P135 = PLOM_UNIFORM64(pv[135]);
    const double P136 = PLOM_UNIFORM64(pv[136]);
    const double P137 = PLOM_UNIFORM64(pv[137]);
    const double P138 = PLOM_UNIFORM64(pv[138]);
    const double P139 = PLOM_UNIFORM64(pv[139]);
    const double P140 = PLOM_UNIFORM64(pv[140]);
    const double P141 = PLOM_UNIFORM64(pv[141]);
    const double P142 = PLOM_UNIFORM64(pv[142]);
    const double P143 = PLOM_UNIFORM64(pv[143]);
    const double P144 = PLOM_UNIFORM64(pv[144]);
    const double P145 = PLOM_UNIFORM64(pv[145]);
    const double P146 = PLOM_UNIFORM64(pv[146]);
    const double P147 = PLOM_UNIFORM64(pv[147]);
    const double P148 = PLOM_UNIFORM64(pv[148]);
    const double P149 = PLOM_UNIFORM64(pv[149]);
    const double P150 = PLOM_UNIFORM64(pv[150]);
    const double P151 = PLOM_UNIFORM64(pv[151]);
    const double P152 = PLOM_UNIFORM64(pv[152]);
    const double P153 = PLOM_UNIFORM64(pv[153]);
    const double P154 = PLOM_UNIFORM64(pv[154]);
    const double P155 = PLOM_UNIFORM64(pv[155]);
    const double P156 = PLOM_UNIFORM64(pv[156]);
    const double P157 = PLOM_UNIFORM64(pv[157]);
    const double P158 = PLOM_UNIFORM64(pv[158]);
    const double P159 = PLOM_UNIFORM64(pv[159]);
    const double P160 = PLOM_UNIFORM64(pv[160]);
    const double P161 = PLOM_UNIFORM64(pv[161]);
    const double P162 = PLOM_UNIFORM64(pv[162]);
    const double P163 = PLOM_UNIFORM64(pv[163]);
    const double P164 = PLOM_UNIFORM64(pv[164]);
    const double P165 = PLOM_UNIFORM64(pv[165]);
    const double P166 = PLOM_UNIFORM64(pv[166]);
    const double P167 = PLOM_UNIFORM64(pv[167]);
    const double P168 = PLOM_UNIFORM64(pv[168]);
    const double P169 = PLOM_UNIFORM64(pv[169]);
    const double P170 = PLOM_UNIFORM64(pv[170]);
    const double P171 = PLOM_UNIFORM64(pv[171]);
    const double P172 = PLOM_UNIFORM64(pv[172]);
    const double P173 = PLOM_UNIFORM64(pv[173]);
    const double P174 = PLOM_UNIFORM64(pv[174]);
    const double P175 = PLOM_UNIFORM64(pv[175]);
    const double P176 = PLOM_UNIFORM64(pv[176]);
    const double P177 = PLOM_UNIFORM64(pv[177]);
    const double P178 = PLOM_UNIFORM64(pv[178]);
    const double P179 = PLOM_UNIFORM64(pv[179]);
    const double C0 = sp ? plom_spline_at(sp, 0, t) : plom_cov_ld(cv + 0);
    const double C1 = sp ? plom_spline_at(sp, 1, t) : plom_cov_ld(cv + 1);
    const double C2 = sp ? plom_spline_at(sp, 2, t) : plom_cov_ld(cv + 2);
    const double C3 = sp ? plom_spline_at(sp, 3, t) : plom_cov_ld(cv + 3);
    const double C4 = sp ? plom_spline_at(sp, 4, t) : plom_cov_ld(cv + 4);
    const double C5 = sp ? plom_spline_at(sp, 5, t) : plom_cov_ld(cv + 5);
    const double C6 = sp ? plom_spline_at(sp, 6, t) : plom_cov_ld(cv + 6);
    const double C7 = sp ? plom_spline_at(sp, 7, t) : plom_cov_ld(cv + 7);
    const double C8 = sp ? plom_spline_at(sp, 8, t) : plom_cov_ld(cv + 8);
    const double C9 = sp ? plom_spline_at(sp, 9, t) : plom_cov_ld(cv + 9);
    const double r0_0 = ((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])) / C0);
    const double r1_0 = P10;
    const double r2_0 = P20;
    const double r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
    const double r0_1 = ((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])) / C1);
    const double r1_1 = P11;
    const double r2_1 = P21;
    const double r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
    const double r0_2 = ((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])) / C2);
    const double r1_2 = P12;
    const double r2_2 = P22;
    const double r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
    const double r0_3 = ((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])) / C3);
    const double r1_3 = P13;
    const double r2_3 = P23;
    const double r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
    const double r0_4 = ((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])) / C4);
    const double r1_4 = P14;
    const double r2_4 = P24;
    const double r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
    const double r0_5 = ((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])) / C5);
    const double r1_5 = P15;
    const double r2_5 = P25;
    const double r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
    const double r0_6 = ((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])) / C6);
    const double r1_6 = P16;
    const double r2_6 = P26;
    const double r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
    const double r0_7 = ((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])) / C7);
    const double r1_7 = P17;
    const double r2_7 = P27;
    const double r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
    const double r0_8 = ((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])) / C8);
    const double r1_8 = P18;
    const double r2_8 = P28;
    const double r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
    const double r0_9 = ((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])) / C9);
    const double r1_9 = P19;
    const double r2_9 = P29;
    const double r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
    f[0] =  - (r0_0*X[0]) + (r3_0);
    f[10] =  - (r1_0*X[10]) + (r0_0*X[0]);
    f[20] =  - (r2_0*X[20]) + (r1_0*X[10]);
    f[1] =  - (r0_1*X[1]) + (r3_1);
    f[11] =  - (r1_1*X[11]) + (r0_1*X[1]);
    f[21] =  - (r2_1*X[21]) + (r1_1*X[11]);
    f[2] =  - (r0_2*X[2]) + (r3_2);
    f[12] =  - (r1_2*X[12]) + (r0_2*X[2]);
    f[22] =  - (r2_2*X[22]) + (r1_2*X[12]);
    f[3] =  - (r0_3*X[3]) + (r3_3);
    f[13] =  - (r1_3*X[13]) + (r0_3*X[3]);
    f[23] =  - (r2_3*X[23]) + (r1_3*X[13]);
    f[4] =  - (r0_4*X[4]) + (r3_4);
    f[14] =  - (r1_4*X[14]) + (r0_4*X[4]);
    f[24] =  - (r2_4*X[24]) + (r1_4*X[14]);
    f[5] =  - (r0_5*X[5]) + (r3_5);
    f[15] =  - (r1_5*X[15]) + (r0_5*X[5]);
    f[25] =  - (r2_5*X[25]) + (r1_5*X[15]);
    f[6] =  - (r0_6*X[6]) + (r3_6);
    f[16] =  - (r1_6*X[16]) + (r0_6*X[6]);
    f[26] =  - (r2_6*X[26]) + (r1_6*X[16]);
    f[7] =  - (r0_7*X[7]) + (r3_7);
    f[17] =  - (r1_7*X[17]) + (r0_7*X[7]);
    f[27] =  - (r2_7*X[27]) + (r1_7*X[17]);
    f[8] =  - (r0_8*X[8]) + (r3_8);
    f[18] =  - (r1_8*X[18]) + (r0_8*X[8]);
    f[28] =  - (r2_8*X[28]) + (r1_8*X[18]);
    f[9] =  - (r0_9*X[9]) + (r3_9);
    f[19] =  - (r1_9*X[19]) + (r0_9*X[9]);
    f[29] =  - (r2_9*X[29]) + (r1_9*X[19]);
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_0*X[10]);
        f[30] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_1*X[11]);
        f[31] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_2*X[12]);
        f[32] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_3*X[13]);
        f[33] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_4*X[14]);
        f[34] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_5*X[15]);
        f[35] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_6*X[16]);
        f[36] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_7*X[17]);
        f[37] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_8*X[18]);
        f[38] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_9*X[19]);
        f[39] = sum_inc;
    }
}

#define PLOM_ODE_FN plom_seir10_ode_integrate
#define PLOM_ODE_RHS plom_seir10_ode_rhs
#define PLOM_ODE_N 40
#define PLOM_ODE_QUAL PLOM_HD
#include "plom_ode_impl.h"

PLOM_HD int plom_seir10_interval_ode_f64(double *X, const double *pv, const double *cv, const plom_spline_set *sp, double t0, double t1, double dt, double eps_abs, double eps_rel)
{
    X[30] = 0.0; /* reset_inc */
    X[31] = 0.0; /* reset_inc */
    X[32] = 0.0; /* reset_inc */
    X[33] = 0.0; /* reset_inc */
    X[34] = 0.0; /* reset_inc */
    X[35] = 0.0; /* reset_inc */
    X[36] = 0.0; /* reset_inc */
    X[37] = 0.0; /* reset_inc */
    X[38] = 0.0; /* reset_inc */
    X[39] = 0.0; /* reset_inc */
    return plom_seir10_ode_integrate(X, pv, 0, cv, sp, t0, t1, dt, eps_abs, eps_rel);
}

PLOM_HD void plom_seir10_interval_f64(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, unsigned int noises_off, const double *cv, int cs, double eps_abs, double eps_rel, const plom_drift_view *dm, const plom_spline_set *sp)
{
    if (noises_off & PLOM_IMPL_ODE) {
        (void)plom_seir10_interval_ode_f64(X, pv, cv, sp, t0, t1, dt, eps_abs, eps_rel);
        return;
    }
    if (noises_off & PLOM_IMPL_PSR) {
        if (noises_off & PLOM_NO_ENV_STO) plom_seir10_interval_psr_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
        else plom_seir10_interval_psr_full_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
        return;
    }
    const unsigned int v = noises_off & (PLOM_NO_DEM_STO | PLOM_NO_ENV_STO);
    if (v == 0u) plom_seir10_interval_full_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
    else if (v == PLOM_NO_DEM_STO) plom_seir10_interval_no_dem_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
    else if (v == PLOM_NO_ENV_STO) plom_seir10_interval_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
    else plom_seir10_interval_no_dem_sto_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, 0);
}

#if defined(__HIPCC__)
static __device__ inline __attribute__((always_inline)) void plom_seir10_time_funcs_f32(float tr, float *out)
{
    (void)tr; (void)out;
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_interval_full_f32(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
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
    const float P20 = PLOM_UNIFORM32(pv[20]);
    const float P21 = PLOM_UNIFORM32(pv[21]);
    const float P22 = PLOM_UNIFORM32(pv[22]);
    const float P23 = PLOM_UNIFORM32(pv[23]);
    const float P24 = PLOM_UNIFORM32(pv[24]);
    const float P25 = PLOM_UNIFORM32(pv[25]);
    const float P26 = PLOM_UNIFORM32(pv[26]);
    const float P27 = PLOM_UNIFORM32(pv[27]);
    const float P28 = PLOM_UNIFORM32(pv[28]);
    const float P29 = PLOM_UNIFORM32(pv[29]);
    const float P30 = PLOM_UNIFORM32(pv[30]);
    const float P31 = PLOM_UNIFORM32(pv[31]);
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
    const float P48 = PLOM_UNIFORM32(pv[48]);
    const float P49 = PLOM_UNIFORM32(pv[49]);
    const float P80 = PLOM_UNIFORM32(pv[80]);
    const float P81 = PLOM_UNIFORM32(pv[81]);
    const float P82 = PLOM_UNIFORM32(pv[82]);
    const float P83 = PLOM_UNIFORM32(pv[83]);
    const float P84 = PLOM_UNIFORM32(pv[84]);
    const float P85 = PLOM_UNIFORM32(pv[85]);
    const float P86 = PLOM_UNIFORM32(pv[86]);
    const float P87 = PLOM_UNIFORM32(pv[87]);
    const float P88 = PLOM_UNIFORM32(pv[88]);
    const float P89 = PLOM_UNIFORM32(pv[89]);
    const float P90 = PLOM_UNIFORM32(pv[90]);
    const float P91 = PLOM_UNIFORM32(pv[91]);
    const float P92 = PLOM_UNIFORM32(pv[92]);
    const float P93 = PLOM_UNIFORM32(pv[93]);
    const float P94 = PLOM_UNIFORM32(pv[94]);
    const float P95 = PLOM_UNIFORM32(pv[95]);
    const float P96 = PLOM_UNIFORM32(pv[96]);
    const float P97 = PLOM_UNIFORM32(pv[97]);
    const float P98 = PLOM_UNIFORM32(pv[98]);
    const float P99 = PLOM_UNIFORM32(pv[99]);
    const float P100 = PLOM_UNIFORM32(pv[100]);
    const float P101 = PLOM_UNIFORM32(pv[101]);
    const float P102 = PLOM_UNIFORM32(pv[102]);
    const float P103 = PLOM_UNIFORM32(pv[103]);
    const float P104 = PLOM_UNIFORM32(pv[104]);
    const float P105 = PLOM_UNIFORM32(pv[105]);
    const float P106 = PLOM_UNIFORM32(pv[106]);
    const float P107 = PLOM_UNIFORM32(pv[107]);
    const float P108 = PLOM_UNIFORM32(pv[108]);
    const float P109 = PLOM_UNIFORM32(pv[109]);
    const float P110 = PLOM_UNIFORM32(pv[110]);
    const float P111 = PLOM_UNIFORM32(pv[111]);
    const float P112 = PLOM_UNIFORM32(pv[112]);
    const float P113 = PLOM_UNIFORM32(pv[113]);
    const float P114 = PLOM_UNIFORM32(pv[114]);
    const float P115 = PLOM_UNIFORM32(pv[115]);
    const float P116 = PLOM_UNIFORM32(pv[116]);
    const float P117 = PLOM_UNIFORM32(pv[117]);
    const float P118 = PLOM_UNIFORM32(pv[118]);
    const float P119 = PLOM_UNIFORM32(pv[119]);
    const float P120 = PLOM_UNIFORM32(pv[120]);
    const float P121 = PLOM_UNIFORM32(pv[121]);
    const float P122 = PLOM_UNIFORM32(pv[122]);
    const float P123 = PLOM_UNIFORM32(pv[123]);
    const float P124 = PLOM_UNIFORM32(pv[124]);
    const float P125 = PLOM_UNIFORM32(pv[125]);
    const float P126 = PLOM_UNIFORM32(pv[126]);
    const float P127 = PLOM_UNIFORM32(pv[127]);
    const float P128 = PLOM_UNIFORM32(pv[128]);
    const float P129 = PLOM_UNIFORM32(pv[129]);
    const float P130 = PLOM_UNIFORM32(pv[130]);
    const float P131 = PLOM_UNIFORM32(pv[131]);
    const float P132 = PLOM_UNIFORM32(pv[132]);
    const float P133 = PLOM_UNIFORM32(pv[133]);
    const float P134 = PLOM_UNIFORM32(pv[134]);
    const float P135 = PLOM_UNIFORM32(pv[135]);
    const float P136 = PLOM_UNIFORM32(pv[136]);
    const float P137 = PLOM_UNIFORM32(pv[137]);
    const float P138 = PLOM_UNIFORM32(pv[138]);
    const float P139 = PLOM_UNIFORM32(pv[139]);
    const float P140 = PLOM_UNIFORM32(pv[140]);
    const float P141 = PLOM_UNIFORM32(pv[141]);
    const float P142 = PLOM_UNIFORM32(pv[142]);
    const float P143 = PLOM_UNIFORM32(pv[143]);
    const float P144 = PLOM_UNIFORM32(pv[144]);
    const float P145 = PLOM_UNIFORM32(pv[145]);
    const float P146 = PLOM_UNIFORM32(pv[146]);
    const float P147 = PLOM_UNIFORM32(pv[147]);
    const float P148 = PLOM_UNIFORM32(pv[148]);
    const float P149 = PLOM_UNIFORM32(pv[149]);
    const float P150 = PLOM_UNIFORM32(pv[150]);
    const float P151 = PLOM_UNIFORM32(pv[151]);
    const float P152 = PLOM_UNIFORM32(pv[152]);
    const float P153 = PLOM_UNIFORM32(pv[153]);
    const float P154 = PLOM_UNIFORM32(pv[154]);
    const float P155 = PLOM_UNIFORM32(pv[155]);
    const float P156 = PLOM_UNIFORM32(pv[156]);
    const float P157 = PLOM_UNIFORM32(pv[157]);
    const float P158 = PLOM_UNIFORM32(pv[158]);
    const float P159 = PLOM_UNIFORM32(pv[159]);
    const float P160 = PLOM_UNIFORM32(pv[160]);
    const float P161 = PLOM_UNIFORM32(pv[161]);
    const float P162 = PLOM_UNIFORM32(pv[162]);
    const float P163 = PLOM_UNIFORM32(pv[163]);
    const float P164 = PLOM_UNIFORM32(pv[164]);
    const float P165 = PLOM_UNIFORM32(pv[165]);
    const float P166 = PLOM_UNIFORM32(pv[166]);
    const float P167 = PLOM_UNIFORM32(pv[167]);
    const float P168 = PLOM_UNIFORM32(pv[168]);
    const float P169 = PLOM_UNIFORM32(pv[169]);
    const float P170 = PLOM_UNIFORM32(pv[170]);
    const float P171 = PLOM_UNIFORM32(pv[171]);
    const float P172 = PLOM_UNIFORM32(pv[172]);
    const float P173 = PLOM_UNIFORM32(pv[173]);
    const float P174 = PLOM_UNIFORM32(pv[174]);
    const float P175 = PLOM_UNIFORM32(pv[175]);
    const float P176 = PLOM_UNIFORM32(pv[176]);
    const float P177 = PLOM_UNIFORM32(pv[177]);
    const float P178 = PLOM_UNIFORM32(pv[178]);
    const float P179 = PLOM_UNIFORM32(pv[179]);
    X[30] = 0.0f; /* reset_inc */
    X[31] = 0.0f; /* reset_inc */
    X[32] = 0.0f; /* reset_inc */
    X[33] = 0.0f; /* reset_inc */
    X[34] = 0.0f; /* reset_inc */
    X[35] = 0.0f; /* reset_inc */
    X[36] = 0.0f; /* reset_inc */
    X[37] = 0.0f; /* reset_inc */
    X[38] = 0.0f; /* reset_inc */
    X[39] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        const float C1 = (float)plom_cov_ld(cvk + 1);
        const float C2 = (float)plom_cov_ld(cvk + 2);
        const float C3 = (float)plom_cov_ld(cvk + 3);
        const float C4 = (float)plom_cov_ld(cvk + 4);
        const float C5 = (float)plom_cov_ld(cvk + 5);
        const float C6 = (float)plom_cov_ld(cvk + 6);
        const float C7 = (float)plom_cov_ld(cvk + 7);
        const float C8 = (float)plom_cov_ld(cvk + 8);
        const float C9 = (float)plom_cov_ld(cvk + 9);
        float s30 = 0.0f;
        float s31 = 0.0f;
        float s32 = 0.0f;
        float s33 = 0.0f;
        float s34 = 0.0f;
        float s35 = 0.0f;
        float s36 = 0.0f;
        float s37 = 0.0f;
        float s38 = 0.0f;
        float s39 = 0.0f;
        const float r0_0 = plom_fdivf((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])), C0);
        const float r1_0 = P10;
        const float r2_0 = P20;
        const float r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        float za0, zb0;
        plom_normal_pair32(ns, &za0, &zb0);
        const float d0_0 = sqdt * za0;
        const float d1_0 = sqdt * zb0;
        float za1, zb1;
        plom_normal_pair32(ns, &za1, &zb1);
        const float d2_0 = sqdt * za1;
        const float d3_0 = sqdt * zb1;
        float za2, zb2;
        plom_normal_pair32(ns, &za2, &zb2);
        const float w0_0 = sqdt * za2;
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r3_0))*dtr + - plom_fsqrtf((r0_0*X[0]))*d0_0+ plom_fsqrtf((r3_0))*d3_0 + - (r0_0*X[0])*P40*w0_0;
        const float f10 = X[10] + ( - (r1_0*X[10]) + (r0_0*X[0]))*dtr + - plom_fsqrtf((r1_0*X[10]))*d1_0+ plom_fsqrtf((r0_0*X[0]))*d0_0 + + (r0_0*X[0])*P40*w0_0;
        const float f20 = X[20] + ( - (r2_0*X[20]) + (r1_0*X[10]))*dtr + - plom_fsqrtf((r2_0*X[20]))*d2_0+ plom_fsqrtf((r1_0*X[10]))*d1_0;
        s30 += ( + (r1_0*X[10]))*dtr + + plom_fsqrtf((r1_0*X[10]))*d1_0;
        X[0] = plom_clipf(f0);
        X[10] = plom_clipf(f10);
        const float r0_1 = plom_fdivf((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])), C1);
        const float r1_1 = P11;
        const float r2_1 = P21;
        const float r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        const float d0_1 = sqdt * zb2;
        float za3, zb3;
        plom_normal_pair32(ns, &za3, &zb3);
        const float d1_1 = sqdt * za3;
        const float d2_1 = sqdt * zb3;
        float za4, zb4;
        plom_normal_pair32(ns, &za4, &zb4);
        const float d3_1 = sqdt * za4;
        const float w0_1 = sqdt * zb4;
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r3_1))*dtr + - plom_fsqrtf((r0_1*X[1]))*d0_1+ plom_fsqrtf((r3_1))*d3_1 + - (r0_1*X[1])*P41*w0_1;
        const float f11 = X[11] + ( - (r1_1*X[11]) + (r0_1*X[1]))*dtr + - plom_fsqrtf((r1_1*X[11]))*d1_1+ plom_fsqrtf((r0_1*X[1]))*d0_1 + + (r0_1*X[1])*P41*w0_1;
        const float f21 = X[21] + ( - (r2_1*X[21]) + (r1_1*X[11]))*dtr + - plom_fsqrtf((r2_1*X[21]))*d2_1+ plom_fsqrtf((r1_1*X[11]))*d1_1;
        s31 += ( + (r1_1*X[11]))*dtr + + plom_fsqrtf((r1_1*X[11]))*d1_1;
        X[1] = plom_clipf(f1);
        X[11] = plom_clipf(f11);
        const float r0_2 = plom_fdivf((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])), C2);
        const float r1_2 = P12;
        const float r2_2 = P22;
        const float r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        float za5, zb5;
        plom_normal_pair32(ns, &za5, &zb5);
        const float d0_2 = sqdt * za5;
        const float d1_2 = sqdt * zb5;
        float za6, zb6;
        plom_normal_pair32(ns, &za6, &zb6);
        const float d2_2 = sqdt * za6;
        const float d3_2 = sqdt * zb6;
        float za7, zb7;
        plom_normal_pair32(ns, &za7, &zb7);
        const float w0_2 = sqdt * za7;
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r3_2))*dtr + - plom_fsqrtf((r0_2*X[2]))*d0_2+ plom_fsqrtf((r3_2))*d3_2 + - (r0_2*X[2])*P42*w0_2;
        const float f12 = X[12] + ( - (r1_2*X[12]) + (r0_2*X[2]))*dtr + - plom_fsqrtf((r1_2*X[12]))*d1_2+ plom_fsqrtf((r0_2*X[2]))*d0_2 + + (r0_2*X[2])*P42*w0_2;
        const float f22 = X[22] + ( - (r2_2*X[22]) + (r1_2*X[12]))*dtr + - plom_fsqrtf((r2_2*X[22]))*d2_2+ plom_fsqrtf((r1_2*X[12]))*d1_2;
        s32 += ( + (r1_2*X[12]))*dtr + + plom_fsqrtf((r1_2*X[12]))*d1_2;
        X[2] = plom_clipf(f2);
        X[12] = plom_clipf(f12);
        const float r0_3 = plom_fdivf((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])), C3);
        const float r1_3 = P13;
        const float r2_3 = P23;
        const float r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        const float d0_3 = sqdt * zb7;
        float za8, zb8;
        plom_normal_pair32(ns, &za8, &zb8);
        const float d1_3 = sqdt * za8;
        const float d2_3 = sqdt * zb8;
        float za9, zb9;
        plom_normal_pair32(ns, &za9, &zb9);
        const float d3_3 = sqdt * za9;
        const float w0_3 = sqdt * zb9;
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r3_3))*dtr + - plom_fsqrtf((r0_3*X[3]))*d0_3+ plom_fsqrtf((r3_3))*d3_3 + - (r0_3*X[3])*P43*w0_3;
        const float f13 = X[13] + ( - (r1_3*X[13]) + (r0_3*X[3]))*dtr + - plom_fsqrtf((r1_3*X[13]))*d1_3+ plom_fsqrtf((r0_3*X[3]))*d0_3 + + (r0_3*X[3])*P43*w0_3;
        const float f23 = X[23] + ( - (r2_3*X[23]) + (r1_3*X[13]))*dtr + - plom_fsqrtf((r2_3*X[23]))*d2_3+ plom_fsqrtf((r1_3*X[13]))*d1_3;
        s33 += ( + (r1_3*X[13]))*dtr + + plom_fsqrtf((r1_3*X[13]))*d1_3;
        X[3] = plom_clipf(f3);
        X[13] = plom_clipf(f13);
        const float r0_4 = plom_fdivf((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])), C4);
        const float r1_4 = P14;
        const float r2_4 = P24;
        const float r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        float za10, zb10;
        plom_normal_pair32(ns, &za10, &zb10);
        const float d0_4 = sqdt * za10;
        const float d1_4 = sqdt * zb10;
        float za11, zb11;
        plom_normal_pair32(ns, &za11, &zb11);
        const float d2_4 = sqdt * za11;
        const float d3_4 = sqdt * zb11;
        float za12, zb12;
        plom_normal_pair32(ns, &za12, &zb12);
        const float w0_4 = sqdt * za12;
        const float f4 = X[4] + ( - (r0_4*X[4]) + (r3_4))*dtr + - plom_fsqrtf((r0_4*X[4]))*d0_4+ plom_fsqrtf((r3_4))*d3_4 + - (r0_4*X[4])*P44*w0_4;
        const float f14 = X[14] + ( - (r1_4*X[14]) + (r0_4*X[4]))*dtr + - plom_fsqrtf((r1_4*X[14]))*d1_4+ plom_fsqrtf((r0_4*X[4]))*d0_4 + + (r0_4*X[4])*P44*w0_4;
        const float f24 = X[24] + ( - (r2_4*X[24]) + (r1_4*X[14]))*dtr + - plom_fsqrtf((r2_4*X[24]))*d2_4+ plom_fsqrtf((r1_4*X[14]))*d1_4;
        s34 += ( + (r1_4*X[14]))*dtr + + plom_fsqrtf((r1_4*X[14]))*d1_4;
        X[4] = plom_clipf(f4);
        X[14] = plom_clipf(f14);
        const float r0_5 = plom_fdivf((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])), C5);
        const float r1_5 = P15;
        const float r2_5 = P25;
        const float r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        const float d0_5 = sqdt * zb12;
        float za13, zb13;
        plom_normal_pair32(ns, &za13, &zb13);
        const float d1_5 = sqdt * za13;
        const float d2_5 = sqdt * zb13;
        float za14, zb14;
        plom_normal_pair32(ns, &za14, &zb14);
        const float d3_5 = sqdt * za14;
        const float w0_5 = sqdt * zb14;
        const float f5 = X[5] + ( - (r0_5*X[5]) + (r3_5))*dtr + - plom_fsqrtf((r0_5*X[5]))*d0_5+ plom_fsqrtf((r3_5))*d3_5 + - (r0_5*X[5])*P45*w0_5;
        const float f15 = X[15] + ( - (r1_5*X[15]) + (r0_5*X[5]))*dtr + - plom_fsqrtf((r1_5*X[15]))*d1_5+ plom_fsqrtf((r0_5*X[5]))*d0_5 + + (r0_5*X[5])*P45*w0_5;
        const float f25 = X[25] + ( - (r2_5*X[25]) + (r1_5*X[15]))*dtr + - plom_fsqrtf((r2_5*X[25]))*d2_5+ plom_fsqrtf((r1_5*X[15]))*d1_5;
        s35 += ( + (r1_5*X[15]))*dtr + + plom_fsqrtf((r1_5*X[15]))*d1_5;
        X[5] = plom_clipf(f5);
        X[15] = plom_clipf(f15);
        const float r0_6 = plom_fdivf((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])), C6);
        const float r1_6 = P16;
        const float r2_6 = P26;
        const float r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        float za15, zb15;
        plom_normal_pair32(ns, &za15, &zb15);
        const float d0_6 = sqdt * za15;
        const float d1_6 = sqdt * zb15;
        float za16, zb16;
        plom_normal_pair32(ns, &za16, &zb16);
        const float d2_6 = sqdt * za16;
        const float d3_6 = sqdt * zb16;
        float za17, zb17;
        plom_normal_pair32(ns, &za17, &zb17);
        const float w0_6 = sqdt * za17;
        const float f6 = X[6] + ( - (r0_6*X[6]) + (r3_6))*dtr + - plom_fsqrtf((r0_6*X[6]))*d0_6+ plom_fsqrtf((r3_6))*d3_6 + - (r0_6*X[6])*P46*w0_6;
        const float f16 = X[16] + ( - (r1_6*X[16]) + (r0_6*X[6]))*dtr + - plom_fsqrtf((r1_6*X[16]))*d1_6+ plom_fsqrtf((r0_6*X[6]))*d0_6 + + (r0_6*X[6])*P46*w0_6;
        const float f26 = X[26] + ( - (r2_6*X[26]) + (r1_6*X[16]))*dtr + - plom_fsqrtf((r2_6*X[26]))*d2_6+ plom_fsqrtf((r1_6*X[16]))*d1_6;
        s36 += ( + (r1_6*X[16]))*dtr + + plom_fsqrtf((r1_6*X[16]))*d1_6;
        X[6] = plom_clipf(f6);
        X[16] = plom_clipf(f16);
        const float r0_7 = plom_fdivf((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])), C7);
        const float r1_7 = P17;
        const float r2_7 = P27;
        const float r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        const float d0_7 = sqdt * zb17;
        float za18, zb18;
        plom_normal_pair32(ns, &za18, &zb18);
        const float d1_7 = sqdt * za18;
        const float d2_7 = sqdt * zb18;
        float za19, zb19;
        plom_normal_pair32(ns, &za19, &zb19);
        const float d3_7 = sqdt * za19;
        const float w0_7 = sqdt * zb19;
        const float f7 = X[7] + ( - (r0_7*X[7]) + (r3_7))*dtr + - plom_fsqrtf((r0_7*X[7]))*d0_7+ plom_fsqrtf((r3_7))*d3_7 + - (r0_7*X[7])*P47*w0_7;
        const float f17 = X[17] + ( - (r1_7*X[17]) + (r0_7*X[7]))*dtr + - plom_fsqrtf((r1_7*X[17]))*d1_7+ plom_fsqrtf((r0_7*X[7]))*d0_7 + + (r0_7*X[7])*P47*w0_7;
        const float f27 = X[27] + ( - (r2_7*X[27]) + (r1_7*X[17]))*dtr + - plom_fsqrtf((r2_7*X[27]))*d2_7+ plom_fsqrtf((r1_7*X[17]))*d1_7;
        s37 += ( + (r1_7*X[17]))*dtr + + plom_fsqrtf((r1_7*X[17]))*d1_7;
        X[7] = plom_clipf(f7);
        X[17] = plom_clipf(f17);
        const float r0_8 = plom_fdivf((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])), C8);
        const float r1_8 = P18;
        const float r2_8 = P28;
        const float r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        float za20, zb20;
        plom_normal_pair32(ns, &za20, &zb20);
        const float d0_8 = sqdt * za20;
        const float d1_8 = sqdt * zb20;
        float za21, zb21;
        plom_normal_pair32(ns, &za21, &zb21);
        const float d2_8 = sqdt * za21;
        const float d3_8 = sqdt * zb21;
        float za22, zb22;
        plom_normal_pair32(ns, &za22, &zb22);
        const float w0_8 = sqdt * za22;
        const float f8 = X[8] + ( - (r0_8*X[8]) + (r3_8))*dtr + - plom_fsqrtf((r0_8*X[8]))*d0_8+ plom_fsqrtf((r3_8))*d3_8 + - (r0_8*X[8])*P48*w0_8;
        const float f18 = X[18] + ( - (r1_8*X[18]) + (r0_8*X[8]))*dtr + - plom_fsqrtf((r1_8*X[18]))*d1_8+ plom_fsqrtf((r0_8*X[8]))*d0_8 + + (r0_8*X[8])*P48*w0_8;
        const float f28 = X[28] + ( - (r2_8*X[28]) + (r1_8*X[18]))*dtr + - plom_fsqrtf((r2_8*X[28]))*d2_8+ plom_fsqrtf((r1_8*X[18]))*d1_8;
        s38 += ( + (r1_8*X[18]))*dtr + + plom_fsqrtf((r1_8*X[18]))*d1_8;
        X[8] = plom_clipf(f8);
        X[18] = plom_clipf(f18);
        const float r0_9 = plom_fdivf((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])), C9);
        const float r1_9 = P19;
        const float r2_9 = P29;
        const float r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        const float d0_9 = sqdt * zb22;
        float za23, zb23;
        plom_normal_pair32(ns, &za23, &zb23);
        const float d1_9 = sqdt * za23;
        const float d2_9 = sqdt * zb23;
        float za24, zb24;
        plom_normal_pair32(ns, &za24, &zb24);
        const float d3_9 = sqdt * za24;
        const float w0_9 = sqdt * zb24;
        const float f9 = X[9] + ( - (r0_9*X[9]) + (r3_9))*dtr + - plom_fsqrtf((r0_9*X[9]))*d0_9+ plom_fsqrtf((r3_9))*d3_9 + - (r0_9*X[9])*P49*w0_9;
        const float f19 = X[19] + ( - (r1_9*X[19]) + (r0_9*X[9]))*dtr + - plom_fsqrtf((r1_9*X[19]))*d1_9+ plom_fsqrtf((r0_9*X[9]))*d0_9 + + (r0_9*X[9])*P49*w0_9;
        const float f29 = X[29] + ( - (r2_9*X[29]) + (r1_9*X[19]))*dtr + - plom_fsqrtf((r2_9*X[29]))*d2_9+ plom_fsqrtf((r1_9*X[19]))*d1_9;
        s39 += ( + (r1_9*X[19]))*dtr + + plom_fsqrtf((r1_9*X[19]))*d1_9;
        X[9] = plom_clipf(f9);
        X[19] = plom_clipf(f19);
        X[20] = plom_clipf(f20);
        X[21] = plom_clipf(f21);
        X[22] = plom_clipf(f22);
        X[23] = plom_clipf(f23);
        X[24] = plom_clipf(f24);
        X[25] = plom_clipf(f25);
        X[26] = plom_clipf(f26);
        X[27] = plom_clipf(f27);
        X[28] = plom_clipf(f28);
        X[29] = plom_clipf(f29);
        const float f30 = X[30] + s30;
        X[30] = plom_clipf(f30);
        const float f31 = X[31] + s31;
        X[31] = plom_clipf(f31);
        const float f32 = X[32] + s32;
        X[32] = plom_clipf(f32);
        const float f33 = X[33] + s33;
        X[33] = plom_clipf(f33);
        const float f34 = X[34] + s34;
        X[34] = plom_clipf(f34);
        const float f35 = X[35] + s35;
        X[35] = plom_clipf(f35);
        const float f36 = X[36] + s36;
        X[36] = plom_clipf(f36);
        const float f37 = X[37] + s37;
        X[37] = plom_clipf(f37);
        const float f38 = X[38] + s38;
        X[38] = plom_clipf(f38);
        const float f39 = X[39] + s39;
        X[39] = plom_clipf(f39);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_interval_no_dem_sto_f32(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
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
    const float P20 = PLOM_UNIFORM32(pv[20]);
    const float P21 = PLOM_UNIFORM32(pv[21]);
    const float P22 = PLOM_UNIFORM32(pv[22]);
    const float P23 = PLOM_UNIFORM32(pv[23]);
    const float P24 = PLOM_UNIFORM32(pv[24]);
    const float P25 = PLOM_UNIFORM32(pv[25]);
    const float P26 = PLOM_UNIFORM32(pv[26]);
    const float P27 = PLOM_UNIFORM32(pv[27]);
    const float P28 = PLOM_UNIFORM32(pv[28]);
    const float P29 = PLOM_UNIFORM32(pv[29]);
    const float P30 = PLOM_UNIFORM32(pv[30]);
    const float P31 = PLOM_UNIFORM32(pv[31]);
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
    const float P48 = PLOM_UNIFORM32(pv[48]);
    const float P49 = PLOM_UNIFORM32(pv[49]);
    const float P80 = PLOM_UNIFORM32(pv[80]);
    const float P81 = PLOM_UNIFORM32(pv[81]);
    const float P82 = PLOM_UNIFORM32(pv[82]);
    const float P83 = PLOM_UNIFORM32(pv[83]);
    const float P84 = PLOM_UNIFORM32(pv[84]);
    const float P85 = PLOM_UNIFORM32(pv[85]);
    const float P86 = PLOM_UNIFORM32(pv[86]);
    const float P87 = PLOM_UNIFORM32(pv[87]);
    const float P88 = PLOM_UNIFORM32(pv[88]);
    const float P89 = PLOM_UNIFORM32(pv[89]);
    const float P90 = PLOM_UNIFORM32(pv[90]);
    const float P91 = PLOM_UNIFORM32(pv[91]);
    const float P92 = PLOM_UNIFORM32(pv[92]);
    const float P93 = PLOM_UNIFORM32(pv[93]);
    const float P94 = PLOM_UNIFORM32(pv[94]);
    const float P95 = PLOM_UNIFORM32(pv[95]);
    const float P96 = PLOM_UNIFORM32(pv[96]);
    const float P97 = PLOM_UNIFORM32(pv[97]);
    const float P98 = PLOM_UNIFORM32(pv[98]);
    const float P99 = PLOM_UNIFORM32(pv[99]);
    const float P100 = PLOM_UNIFORM32(pv[100]);
    const float P101 = PLOM_UNIFORM32(pv[101]);
    const float P102 = PLOM_UNIFORM32(pv[102]);
    const float P103 = PLOM_UNIFORM32(pv[103]);
    const float P104 = PLOM_UNIFORM32(pv[104]);
    const float P105 = PLOM_UNIFORM32(pv[105]);
    const float P106 = PLOM_UNIFORM32(pv[106]);
    const float P107 = PLOM_UNIFORM32(pv[107]);
    const float P108 = PLOM_UNIFORM32(pv[108]);
    const float P109 = PLOM_UNIFORM32(pv[109]);
    const float P110 = PLOM_UNIFORM32(pv[110]);
    const float P111 = PLOM_UNIFORM32(pv[111]);
    const float P112 = PLOM_UNIFORM32(pv[112]);
    const float P113 = PLOM_UNIFORM32(pv[113]);
    const float P114 = PLOM_UNIFORM32(pv[114]);
    const float P115 = PLOM_UNIFORM32(pv[115]);
    const float P116 = PLOM_UNIFORM32(pv[116]);
    const float P117 = PLOM_UNIFORM32(pv[117]);
    const float P118 = PLOM_UNIFORM32(pv[118]);
    const float P119 = PLOM_UNIFORM32(pv[119]);
    const float P120 = PLOM_UNIFORM32(pv[120]);
    const float P121 = PLOM_UNIFORM32(pv[121]);
    const float P122 = PLOM_UNIFORM32(pv[122]);
    const float P123 = PLOM_UNIFORM32(pv[123]);
    const float P124 = PLOM_UNIFORM32(pv[124]);
    const float P125 = PLOM_UNIFORM32(pv[125]);
    const float P126 = PLOM_UNIFORM32(pv[126]);
    const float P127 = PLOM_UNIFORM32(pv[127]);
    const float P128 = PLOM_UNIFORM32(pv[128]);
    const float P129 = PLOM_UNIFORM32(pv[129]);
    const float P130 = PLOM_UNIFORM32(pv[130]);
    const float P131 = PLOM_UNIFORM32(pv[131]);
    const float P132 = PLOM_UNIFORM32(pv[132]);
    const float P133 = PLOM_UNIFORM32(pv[133]);
    const float P134 = PLOM_UNIFORM32(pv[134]);
    const float P135 = PLOM_UNIFORM32(pv[135]);
    const float P136 = PLOM_UNIFORM32(pv[136]);
    const float P137 = PLOM_UNIFORM32(pv[137]);
    const float P138 = PLOM_UNIFORM32(pv[138]);
    const float P139 = PLOM_UNIFORM32(pv[139]);
    const float P140 = PLOM_UNIFORM32(pv[140]);
    const float P141 = PLOM_UNIFORM32(pv[141]);
    const float P142 = PLOM_UNIFORM32(pv[142]);
    const float P143 = PLOM_UNIFORM32(pv[143]);
    const float P144 = PLOM_UNIFORM32(pv[144]);
    const float P145 = PLOM_UNIFORM32(pv[145]);
    const float P146 = PLOM_UNIFORM32(pv[146]);
    const float P147 = PLOM_UNIFORM32(pv[147]);
    const float P148 = PLOM_UNIFORM32(pv[148]);
    const float P149 = PLOM_UNIFORM32(pv[149]);
    const float P150 = PLOM_UNIFORM32(pv[150]);
    const float P151 = PLOM_UNIFORM32(pv[151]);
    const float P152 = PLOM_UNIFORM32(pv[152]);
    const float P153 = PLOM_UNIFORM32(pv[153]);
    const float P154 = PLOM_UNIFORM32(pv[154]);
    const float P155 = PLOM_UNIFORM32(pv[155]);
    const float P156 = PLOM_UNIFORM32(pv[156]);
    const float P157 = PLOM_UNIFORM32(pv[157]);
    const float P158 = PLOM_UNIFORM32(pv[158]);
    const float P159 = PLOM_UNIFORM32(pv[159]);
    const float P160 = PLOM_UNIFORM32(pv[160]);
    const float P161 = PLOM_UNIFORM32(pv[161]);
    const float P162 = PLOM_UNIFORM32(pv[162]);
    const float P163 = PLOM_UNIFORM32(pv[163]);
    const float P164 = PLOM_UNIFORM32(pv[164]);
    const float P165 = PLOM_UNIFORM32(pv[165]);
    const float P166 = PLOM_UNIFORM32(pv[166]);
    const float P167 = PLOM_UNIFORM32(pv[167]);
    const float P168 = PLOM_UNIFORM32(pv[168]);
    const float P169 = PLOM_UNIFORM32(pv[169]);
    const float P170 = PLOM_UNIFORM32(pv[170]);
    const float P171 = PLOM_UNIFORM32(pv[171]);
    const float P172 = PLOM_UNIFORM32(pv[172]);
    const float P173 = PLOM_UNIFORM32(pv[173]);
    const float P174 = PLOM_UNIFORM32(pv[174]);
    const float P175 = PLOM_UNIFORM32(pv[175]);
    const float P176 = PLOM_UNIFORM32(pv[176]);
    const float P177 = PLOM_UNIFORM32(pv[177]);
    const float P178 = PLOM_UNIFORM32(pv[178]);
    const float P179 = PLOM_UNIFORM32(pv[179]);
    X[30] = 0.0f; /* reset_inc */
    X[31] = 0.0f; /* reset_inc */
    X[32] = 0.0f; /* reset_inc */
    X[33] = 0.0f; /* reset_inc */
    X[34] = 0.0f; /* reset_inc */
    X[35] = 0.0f; /* reset_inc */
    X[36] = 0.0f; /* reset_inc */
    X[37] = 0.0f; /* reset_inc */
    X[38] = 0.0f; /* reset_inc */
    X[39] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        const float C1 = (float)plom_cov_ld(cvk + 1);
        const float C2 = (float)plom_cov_ld(cvk + 2);
        const float C3 = (float)plom_cov_ld(cvk + 3);
        const float C4 = (float)plom_cov_ld(cvk + 4);
        const float C5 = (float)plom_cov_ld(cvk + 5);
        const float C6 = (float)plom_cov_ld(cvk + 6);
        const float C7 = (float)plom_cov_ld(cvk + 7);
        const float C8 = (float)plom_cov_ld(cvk + 8);
        const float C9 = (float)plom_cov_ld(cvk + 9);
        float s30 = 0.0f;
        float s31 = 0.0f;
        float s32 = 0.0f;
        float s33 = 0.0f;
        float s34 = 0.0f;
        float s35 = 0.0f;
        float s36 = 0.0f;
        float s37 = 0.0f;
        float s38 = 0.0f;
        float s39 = 0.0f;
        const float r0_0 = plom_fdivf((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])), C0);
        const float r1_0 = P10;
        const float r2_0 = P20;
        const float r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        float za0, zb0;
        plom_normal_pair32(ns, &za0, &zb0);
        const float w0_0 = sqdt * za0;
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r3_0))*dtr + - (r0_0*X[0])*P40*w0_0;
        const float f10 = X[10] + ( - (r1_0*X[10]) + (r0_0*X[0]))*dtr + + (r0_0*X[0])*P40*w0_0;
        const float f20 = X[20] + ( - (r2_0*X[20]) + (r1_0*X[10]))*dtr;
        s30 += ( + (r1_0*X[10]))*dtr;
        X[0] = plom_clipf(f0);
        X[10] = plom_clipf(f10);
        const float r0_1 = plom_fdivf((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])), C1);
        const float r1_1 = P11;
        const float r2_1 = P21;
        const float r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        const float w0_1 = sqdt * zb0;
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r3_1))*dtr + - (r0_1*X[1])*P41*w0_1;
        const float f11 = X[11] + ( - (r1_1*X[11]) + (r0_1*X[1]))*dtr + + (r0_1*X[1])*P41*w0_1;
        const float f21 = X[21] + ( - (r2_1*X[21]) + (r1_1*X[11]))*dtr;
        s31 += ( + (r1_1*X[11]))*dtr;
        X[1] = plom_clipf(f1);
        X[11] = plom_clipf(f11);
        const float r0_2 = plom_fdivf((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])), C2);
        const float r1_2 = P12;
        const float r2_2 = P22;
        const float r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        float za1, zb1;
        plom_normal_pair32(ns, &za1, &zb1);
        const float w0_2 = sqdt * za1;
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r3_2))*dtr + - (r0_2*X[2])*P42*w0_2;
        const float f12 = X[12] + ( - (r1_2*X[12]) + (r0_2*X[2]))*dtr + + (r0_2*X[2])*P42*w0_2;
        const float f22 = X[22] + ( - (r2_2*X[22]) + (r1_2*X[12]))*dtr;
        s32 += ( + (r1_2*X[12]))*dtr;
        X[2] = plom_clipf(f2);
        X[12] = plom_clipf(f12);
        const float r0_3 = plom_fdivf((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])), C3);
        const float r1_3 = P13;
        const float r2_3 = P23;
        const float r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        const float w0_3 = sqdt * zb1;
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r3_3))*dtr + - (r0_3*X[3])*P43*w0_3;
        const float f13 = X[13] + ( - (r1_3*X[13]) + (r0_3*X[3]))*dtr + + (r0_3*X[3])*P43*w0_3;
        const float f23 = X[23] + ( - (r2_3*X[23]) + (r1_3*X[13]))*dtr;
        s33 += ( + (r1_3*X[13]))*dtr;
        X[3] = plom_clipf(f3);
        X[13] = plom_clipf(f13);
        const float r0_4 = plom_fdivf((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])), C4);
        const float r1_4 = P14;
        const float r2_4 = P24;
        const float r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        float za2, zb2;
        plom_normal_pair32(ns, &za2, &zb2);
        const float w0_4 = sqdt * za2;
        const float f4 = X[4] + ( - (r0_4*X[4]) + (r3_4))*dtr + - (r0_4*X[4])*P44*w0_4;
        const float f14 = X[14] + ( - (r1_4*X[14]) + (r0_4*X[4]))*dtr + + (r0_4*X[4])*P44*w0_4;
        const float f24 = X[24] + ( - (r2_4*X[24]) + (r1_4*X[14]))*dtr;
        s34 += ( + (r1_4*X[14]))*dtr;
        X[4] = plom_clipf(f4);
        X[14] = plom_clipf(f14);
        const float r0_5 = plom_fdivf((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])), C5);
        const float r1_5 = P15;
        const float r2_5 = P25;
        const float r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        const float w0_5 = sqdt * zb2;
        const float f5 = X[5] + ( - (r0_5*X[5]) + (r3_5))*dtr + - (r0_5*X[5])*P45*w0_5;
        const float f15 = X[15] + ( - (r1_5*X[15]) + (r0_5*X[5]))*dtr + + (r0_5*X[5])*P45*w0_5;
        const float f25 = X[25] + ( - (r2_5*X[25]) + (r1_5*X[15]))*dtr;
        s35 += ( + (r1_5*X[15]))*dtr;
        X[5] = plom_clipf(f5);
        X[15] = plom_clipf(f15);
        const float r0_6 = plom_fdivf((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])), C6);
        const float r1_6 = P16;
        const float r2_6 = P26;
        const float r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        float za3, zb3;
        plom_normal_pair32(ns, &za3, &zb3);
        const float w0_6 = sqdt * za3;
        const float f6 = X[6] + ( - (r0_6*X[6]) + (r3_6))*dtr + - (r0_6*X[6])*P46*w0_6;
        const float f16 = X[16] + ( - (r1_6*X[16]) + (r0_6*X[6]))*dtr + + (r0_6*X[6])*P46*w0_6;
        const float f26 = X[26] + ( - (r2_6*X[26]) + (r1_6*X[16]))*dtr;
        s36 += ( + (r1_6*X[16]))*dtr;
        X[6] = plom_clipf(f6);
        X[16] = plom_clipf(f16);
        const float r0_7 = plom_fdivf((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])), C7);
        const float r1_7 = P17;
        const float r2_7 = P27;
        const float r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        const float w0_7 = sqdt * zb3;
        const float f7 = X[7] + ( - (r0_7*X[7]) + (r3_7))*dtr + - (r0_7*X[7])*P47*w0_7;
        const float f17 = X[17] + ( - (r1_7*X[17]) + (r0_7*X[7]))*dtr + + (r0_7*X[7])*P47*w0_7;
        const float f27 = X[27] + ( - (r2_7*X[27]) + (r1_7*X[17]))*dtr;
        s37 += ( + (r1_7*X[17]))*dtr;
        X[7] = plom_clipf(f7);
        X[17] = plom_clipf(f17);
        const float r0_8 = plom_fdivf((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])), C8);
        const float r1_8 = P18;
        const float r2_8 = P28;
        const float r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        float za4, zb4;
        plom_normal_pair32(ns, &za4, &zb4);
        const float w0_8 = sqdt * za4;
        const float f8 = X[8] + ( - (r0_8*X[8]) + (r3_8))*dtr + - (r0_8*X[8])*P48*w0_8;
        const float f18 = X[18] + ( - (r1_8*X[18]) + (r0_8*X[8]))*dtr + + (r0_8*X[8])*P48*w0_8;
        const float f28 = X[28] + ( - (r2_8*X[28]) + (r1_8*X[18]))*dtr;
        s38 += ( + (r1_8*X[18]))*dtr;
        X[8] = plom_clipf(f8);
        X[18] = plom_clipf(f18);
        const float r0_9 = plom_fdivf((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])), C9);
        const float r1_9 = P19;
        const float r2_9 = P29;
        const float r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        const float w0_9 = sqdt * zb4;
        const float f9 = X[9] + ( - (r0_9*X[9]) + (r3_9))*dtr + - (r0_9*X[9])*P49*w0_9;
        const float f19 = X[19] + ( - (r1_9*X[19]) + (r0_9*X[9]))*dtr + + (r0_9*X[9])*P49*w0_9;
        const float f29 = X[29] + ( - (r2_9*X[29]) + (r1_9*X[19]))*dtr;
        s39 += ( + (r1_9*X[19]))*dtr;
        X[9] = plom_clipf(f9);
        X[19] = plom_clipf(f19);
        X[20] = plom_clipf(f20);
        X[21] = plom_clipf(f21);
        X[22] = plom_clipf(f22);
        X[23] = plom_clipf(f23);
        X[24] = plom_clipf(f24);
        X[25] = plom_clipf(f25);
        X[26] = plom_clipf(f26);
        X[27] = plom_clipf(f27);
        X[28] = plom_clipf(f28);
        X[29] = plom_clipf(f29);
        const float f30 = X[30] + s30;
        X[30] = plom_clipf(f30);
        const float f31 = X[31] + s31;
        X[31] = plom_clipf(f31);
        const float f32 = X[32] + s32;
        X[32] = plom_clipf(f32);
        const float f33 = X[33] + s33;
        X[33] = plom_clipf(f33);
        const float f34 = X[34] + s34;
        X[34] = plom_clipf(f34);
        const float f35 = X[35] + s35;
        X[35] = plom_clipf(f35);
        const float f36 = X[36] + s36;
        X[36] = plom_clipf(f36);
        const float f37 = X[37] + s37;
        X[37] = plom_clipf(f37);
        const float f38 = X[38] + s38;
        X[38] = plom_clipf(f38);
        const float f39 = X[39] + s39;
        X[39] = plom_clipf(f39);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_interval_no_env_sto_f32(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
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
    const float P20 = PLOM_UNIFORM32(pv[20]);
    const float P21 = PLOM_UNIFORM32(pv[21]);
    const float P22 = PLOM_UNIFORM32(pv[22]);
    const float P23 = PLOM_UNIFORM32(pv[23]);
    const float P24 = PLOM_UNIFORM32(pv[24]);
    const float P25 = PLOM_UNIFORM32(pv[25]);
    const float P26 = PLOM_UNIFORM32(pv[26]);
    const float P27 = PLOM_UNIFORM32(pv[27]);
    const float P28 = PLOM_UNIFORM32(pv[28]);
    const float P29 = PLOM_UNIFORM32(pv[29]);
    const float P30 = PLOM_UNIFORM32(pv[30]);
    const float P31 = PLOM_UNIFORM32(pv[31]);
    const float P32 = PLOM_UNIFORM32(pv[32]);
    const float P33 = PLOM_UNIFORM32(pv[33]);
    const float P34 = PLOM_UNIFORM32(pv[34]);
    const float P35 = PLOM_UNIFORM32(pv[35]);
    const float P36 = PLOM_UNIFORM32(pv[36]);
    const float P37 = PLOM_UNIFORM32(pv[37]);
    const float P38 = PLOM_UNIFORM32(pv[38]);
    const float P39 = PLOM_UNIFORM32(pv[39]);
    const float P80 = PLOM_UNIFORM32(pv[80]);
    const float P81 = PLOM_UNIFORM32(pv[81]);
    const float P82 = PLOM_UNIFORM32(pv[82]);
    const float P83 = PLOM_UNIFORM32(pv[83]);
    const float P84 = PLOM_UNIFORM32(pv[84]);
    const float P85 = PLOM_UNIFORM32(pv[85]);
    const float P86 = PLOM_UNIFORM32(pv[86]);
    const float P87 = PLOM_UNIFORM32(pv[87]);
    const float P88 = PLOM_UNIFORM32(pv[88]);
    const float P89 = PLOM_UNIFORM32(pv[89]);
    const float P90 = PLOM_UNIFORM32(pv[90]);
    const float P91 = PLOM_UNIFORM32(pv[91]);
    const float P92 = PLOM_UNIFORM32(pv[92]);
    const float P93 = PLOM_UNIFORM32(pv[93]);
    const float P94 = PLOM_UNIFORM32(pv[94]);
    const float P95 = PLOM_UNIFORM32(pv[95]);
    const float P96 = PLOM_UNIFORM32(pv[96]);
    const float P97 = PLOM_UNIFORM32(pv[97]);
    const float P98 = PLOM_UNIFORM32(pv[98]);
    const float P99 = PLOM_UNIFORM32(pv[99]);
    const float P100 = PLOM_UNIFORM32(pv[100]);
    const float P101 = PLOM_UNIFORM32(pv[101]);
    const float P102 = PLOM_UNIFORM32(pv[102]);
    const float P103 = PLOM_UNIFORM32(pv[103]);
    const float P104 = PLOM_UNIFORM32(pv[104]);
    const float P105 = PLOM_UNIFORM32(pv[105]);
    const float P106 = PLOM_UNIFORM32(pv[106]);
    const float P107 = PLOM_UNIFORM32(pv[107]);
    const float P108 = PLOM_UNIFORM32(pv[108]);
    const float P109 = PLOM_UNIFORM32(pv[109]);
    const float P110 = PLOM_UNIFORM32(pv[110]);
    const float P111 = PLOM_UNIFORM32(pv[111]);
    const float P112 = PLOM_UNIFORM32(pv[112]);
    const float P113 = PLOM_UNIFORM32(pv[113]);
    const float P114 = PLOM_UNIFORM32(pv[114]);
    const float P115 = PLOM_UNIFORM32(pv[115]);
    const float P116 = PLOM_UNIFORM32(pv[116]);
    const float P117 = PLOM_UNIFORM32(pv[117]);
    const float P118 = PLOM_UNIFORM32(pv[118]);
    const float P119 = PLOM_UNIFORM32(pv[119]);
    const float P120 = PLOM_UNIFORM32(pv[120]);
    const float P121 = PLOM_UNIFORM32(pv[121]);
    const float P122 = PLOM_UNIFORM32(pv[122]);
    const float P123 = PLOM_UNIFORM32(pv[123]);
    const float P124 = PLOM_UNIFORM32(pv[124]);
    const float P125 = PLOM_UNIFORM32(pv[125]);
    const float P126 = PLOM_UNIFORM32(pv[126]);
    const float P127 = PLOM_UNIFORM32(pv[127]);
    const float P128 = PLOM_UNIFORM32(pv[128]);
    const float P129 = PLOM_UNIFORM32(pv[129]);
    const float P130 = PLOM_UNIFORM32(pv[130]);
    const float P131 = PLOM_UNIFORM32(pv[131]);
    const float P132 = PLOM_UNIFORM32(pv[132]);
    const float P133 = PLOM_UNIFORM32(pv[133]);
    const float P134 = PLOM_UNIFORM32(pv[134]);
    const float P135 = PLOM_UNIFORM32(pv[135]);
    const float P136 = PLOM_UNIFORM32(pv[136]);
    const float P137 = PLOM_UNIFORM32(pv[137]);
    const float P138 = PLOM_UNIFORM32(pv[138]);
    const float P139 = PLOM_UNIFORM32(pv[139]);
    const float P140 = PLOM_UNIFORM32(pv[140]);
    const float P141 = PLOM_UNIFORM32(pv[141]);
    const float P142 = PLOM_UNIFORM32(pv[142]);
    const float P143 = PLOM_UNIFORM32(pv[143]);
    const float P144 = PLOM_UNIFORM32(pv[144]);
    const float P145 = PLOM_UNIFORM32(pv[145]);
    const float P146 = PLOM_UNIFORM32(pv[146]);
    const float P147 = PLOM_UNIFORM32(pv[147]);
    const float P148 = PLOM_UNIFORM32(pv[148]);
    const float P149 = PLOM_UNIFORM32(pv[149]);
    const float P150 = PLOM_UNIFORM32(pv[150]);
    const float P151 = PLOM_UNIFORM32(pv[151]);
    const float P152 = PLOM_UNIFORM32(pv[152]);
    const float P153 = PLOM_UNIFORM32(pv[153]);
    const float P154 = PLOM_UNIFORM32(pv[154]);
    const float P155 = PLOM_UNIFORM32(pv[155]);
    const float P156 = PLOM_UNIFORM32(pv[156]);
    const float P157 = PLOM_UNIFORM32(pv[157]);
    const float P158 = PLOM_UNIFORM32(pv[158]);
    const float P159 = PLOM_UNIFORM32(pv[159]);
    const float P160 = PLOM_UNIFORM32(pv[160]);
    const float P161 = PLOM_UNIFORM32(pv[161]);
    const float P162 = PLOM_UNIFORM32(pv[162]);
    const float P163 = PLOM_UNIFORM32(pv[163]);
    const float P164 = PLOM_UNIFORM32(pv[164]);
    const float P165 = PLOM_UNIFORM32(pv[165]);
    const float P166 = PLOM_UNIFORM32(pv[166]);
    const float P167 = PLOM_UNIFORM32(pv[167]);
    const float P168 = PLOM_UNIFORM32(pv[168]);
    const float P169 = PLOM_UNIFORM32(pv[169]);
    const float P170 = PLOM_UNIFORM32(pv[170]);
    const float P171 = PLOM_UNIFORM32(pv[171]);
    const float P172 = PLOM_UNIFORM32(pv[172]);
    const float P173 = PLOM_UNIFORM32(pv[173]);
    const float P174 = PLOM_UNIFORM32(pv[174]);
    const float P175 = PLOM_UNIFORM32(pv[175]);
    const float P176 = PLOM_UNIFORM32(pv[176]);
    const float P177 = PLOM_UNIFORM32(pv[177]);
    const float P178 = PLOM_UNIFORM32(pv[178]);
    const float P179 = PLOM_UNIFORM32(pv[179]);
    X[30] = 0.0f; /* reset_inc */
    X[31] = 0.0f; /* reset_inc */
    X[32] = 0.0f; /* reset_inc */
    X[33] = 0.0f; /* reset_inc */
    X[34] = 0.0f; /* reset_inc */
    X[35] = 0.0f; /* reset_inc */
    X[36] = 0.0f; /* reset_inc */
    X[37] = 0.0f; /* reset_inc */
    X[38] = 0.0f; /* reset_inc */
    X[39] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        const float C1 = (float)plom_cov_ld(cvk + 1);
        const float C2 = (float)plom_cov_ld(cvk + 2);
        const float C3 = (float)plom_cov_ld(cvk + 3);
        const float C4 = (float)plom_cov_ld(cvk + 4);
        const float C5 = (float)plom_cov_ld(cvk + 5);
        const float C6 = (float)plom_cov_ld(cvk + 6);
        const float C7 = (float)plom_cov_ld(cvk + 7);
        const float C8 = (float)plom_cov_ld(cvk + 8);
        const float C9 = (float)plom_cov_ld(cvk + 9);
        float s30 = 0.0f;
        float s31 = 0.0f;
        float s32 = 0.0f;
        float s33 = 0.0f;
        float s34 = 0.0f;
        float s35 = 0.0f;
        float s36 = 0.0f;
        float s37 = 0.0f;
        float s38 = 0.0f;
        float s39 = 0.0f;
        const float r0_0 = plom_fdivf((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])), C0);
        const float r1_0 = P10;
        const float r2_0 = P20;
        const float r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        float za0, zb0;
        plom_normal_pair32(ns, &za0, &zb0);
        const float d0_0 = sqdt * za0;
        const float d1_0 = sqdt * zb0;
        float za1, zb1;
        plom_normal_pair32(ns, &za1, &zb1);
        const float d2_0 = sqdt * za1;
        const float d3_0 = sqdt * zb1;
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r3_0))*dtr + - plom_fsqrtf((r0_0*X[0]))*d0_0+ plom_fsqrtf((r3_0))*d3_0;
        const float f10 = X[10] + ( - (r1_0*X[10]) + (r0_0*X[0]))*dtr + - plom_fsqrtf((r1_0*X[10]))*d1_0+ plom_fsqrtf((r0_0*X[0]))*d0_0;
        const float f20 = X[20] + ( - (r2_0*X[20]) + (r1_0*X[10]))*dtr + - plom_fsqrtf((r2_0*X[20]))*d2_0+ plom_fsqrtf((r1_0*X[10]))*d1_0;
        s30 += ( + (r1_0*X[10]))*dtr + + plom_fsqrtf((r1_0*X[10]))*d1_0;
        X[0] = plom_clipf(f0);
        X[10] = plom_clipf(f10);
        const float r0_1 = plom_fdivf((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])), C1);
        const float r1_1 = P11;
        const float r2_1 = P21;
        const float r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        float za2, zb2;
        plom_normal_pair32(ns, &za2, &zb2);
        const float d0_1 = sqdt * za2;
        const float d1_1 = sqdt * zb2;
        float za3, zb3;
        plom_normal_pair32(ns, &za3, &zb3);
        const float d2_1 = sqdt * za3;
        const float d3_1 = sqdt * zb3;
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r3_1))*dtr + - plom_fsqrtf((r0_1*X[1]))*d0_1+ plom_fsqrtf((r3_1))*d3_1;
        const float f11 = X[11] + ( - (r1_1*X[11]) + (r0_1*X[1]))*dtr + - plom_fsqrtf((r1_1*X[11]))*d1_1+ plom_fsqrtf((r0_1*X[1]))*d0_1;
        const float f21 = X[21] + ( - (r2_1*X[21]) + (r1_1*X[11]))*dtr + - plom_fsqrtf((r2_1*X[21]))*d2_1+ plom_fsqrtf((r1_1*X[11]))*d1_1;
        s31 += ( + (r1_1*X[11]))*dtr + + plom_fsqrtf((r1_1*X[11]))*d1_1;
        X[1] = plom_clipf(f1);
        X[11] = plom_clipf(f11);
        const float r0_2 = plom_fdivf((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])), C2);
        const float r1_2 = P12;
        const float r2_2 = P22;
        const float r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        float za4, zb4;
        plom_normal_pair32(ns, &za4, &zb4);
        const float d0_2 = sqdt * za4;
        const float d1_2 = sqdt * zb4;
        float za5, zb5;
        plom_normal_pair32(ns, &za5, &zb5);
        const float d2_2 = sqdt * za5;
        const float d3_2 = sqdt * zb5;
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r3_2))*dtr + - plom_fsqrtf((r0_2*X[2]))*d0_2+ plom_fsqrtf((r3_2))*d3_2;
        const float f12 = X[12] + ( - (r1_2*X[12]) + (r0_2*X[2]))*dtr + - plom_fsqrtf((r1_2*X[12]))*d1_2+ plom_fsqrtf((r0_2*X[2]))*d0_2;
        const float f22 = X[22] + ( - (r2_2*X[22]) + (r1_2*X[12]))*dtr + - plom_fsqrtf((r2_2*X[22]))*d2_2+ plom_fsqrtf((r1_2*X[12]))*d1_2;
        s32 += ( + (r1_2*X[12]))*dtr + + plom_fsqrtf((r1_2*X[12]))*d1_2;
        X[2] = plom_clipf(f2);
        X[12] = plom_clipf(f12);
        const float r0_3 = plom_fdivf((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])), C3);
        const float r1_3 = P13;
        const float r2_3 = P23;
        const float r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        float za6, zb6;
        plom_normal_pair32(ns, &za6, &zb6);
        const float d0_3 = sqdt * za6;
        const float d1_3 = sqdt * zb6;
        float za7, zb7;
        plom_normal_pair32(ns, &za7, &zb7);
        const float d2_3 = sqdt * za7;
        const float d3_3 = sqdt * zb7;
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r3_3))*dtr + - plom_fsqrtf((r0_3*X[3]))*d0_3+ plom_fsqrtf((r3_3))*d3_3;
        const float f13 = X[13] + ( - (r1_3*X[13]) + (r0_3*X[3]))*dtr + - plom_fsqrtf((r1_3*X[13]))*d1_3+ plom_fsqrtf((r0_3*X[3]))*d0_3;
        const float f23 = X[23] + ( - (r2_3*X[23]) + (r1_3*X[13]))*dtr + - plom_fsqrtf((r2_3*X[23]))*d2_3+ plom_fsqrtf((r1_3*X[13]))*d1_3;
        s33 += ( + (r1_3*X[13]))*dtr + + plom_fsqrtf((r1_3*X[13]))*d1_3;
        X[3] = plom_clipf(f3);
        X[13] = plom_clipf(f13);
        const float r0_4 = plom_fdivf((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])), C4);
        const float r1_4 = P14;
        const float r2_4 = P24;
        const float r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        float za8, zb8;
        plom_normal_pair32(ns, &za8, &zb8);
        const float d0_4 = sqdt * za8;
        const float d1_4 = sqdt * zb8;
        float za9, zb9;
        plom_normal_pair32(ns, &za9, &zb9);
        const float d2_4 = sqdt * za9;
        const float d3_4 = sqdt * zb9;
        const float f4 = X[4] + ( - (r0_4*X[4]) + (r3_4))*dtr + - plom_fsqrtf((r0_4*X[4]))*d0_4+ plom_fsqrtf((r3_4))*d3_4;
        const float f14 = X[14] + ( - (r1_4*X[14]) + (r0_4*X[4]))*dtr + - plom_fsqrtf((r1_4*X[14]))*d1_4+ plom_fsqrtf((r0_4*X[4]))*d0_4;
        const float f24 = X[24] + ( - (r2_4*X[24]) + (r1_4*X[14]))*dtr + - plom_fsqrtf((r2_4*X[24]))*d2_4+ plom_fsqrtf((r1_4*X[14]))*d1_4;
        s34 += ( + (r1_4*X[14]))*dtr + + plom_fsqrtf((r1_4*X[14]))*d1_4;
        X[4] = plom_clipf(f4);
        X[14] = plom_clipf(f14);
        const float r0_5 = plom_fdivf((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])), C5);
        const float r1_5 = P15;
        const float r2_5 = P25;
        const float r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        float za10, zb10;
        plom_normal_pair32(ns, &za10, &zb10);
        const float d0_5 = sqdt * za10;
        const float d1_5 = sqdt * zb10;
        float za11, zb11;
        plom_normal_pair32(ns, &za11, &zb11);
        const float d2_5 = sqdt * za11;
        const float d3_5 = sqdt * zb11;
        const float f5 = X[5] + ( - (r0_5*X[5]) + (r3_5))*dtr + - plom_fsqrtf((r0_5*X[5]))*d0_5+ plom_fsqrtf((r3_5))*d3_5;
        const float f15 = X[15] + ( - (r1_5*X[15]) + (r0_5*X[5]))*dtr + - plom_fsqrtf((r1_5*X[15]))*d1_5+ plom_fsqrtf((r0_5*X[5]))*d0_5;
        const float f25 = X[25] + ( - (r2_5*X[25]) + (r1_5*X[15]))*dtr + - plom_fsqrtf((r2_5*X[25]))*d2_5+ plom_fsqrtf((r1_5*X[15]))*d1_5;
        s35 += ( + (r1_5*X[15]))*dtr + + plom_fsqrtf((r1_5*X[15]))*d1_5;
        X[5] = plom_clipf(f5);
        X[15] = plom_clipf(f15);
        const float r0_6 = plom_fdivf((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])), C6);
        const float r1_6 = P16;
        const float r2_6 = P26;
        const float r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        float za12, zb12;
        plom_normal_pair32(ns, &za12, &zb12);
        const float d0_6 = sqdt * za12;
        const float d1_6 = sqdt * zb12;
        float za13, zb13;
        plom_normal_pair32(ns, &za13, &zb13);
        const float d2_6 = sqdt * za13;
        const float d3_6 = sqdt * zb13;
        const float f6 = X[6] + ( - (r0_6*X[6]) + (r3_6))*dtr + - plom_fsqrtf((r0_6*X[6]))*d0_6+ plom_fsqrtf((r3_6))*d3_6;
        const float f16 = X[16] + ( - (r1_6*X[16]) + (r0_6*X[6]))*dtr + - plom_fsqrtf((r1_6*X[16]))*d1_6+ plom_fsqrtf((r0_6*X[6]))*d0_6;
        const float f26 = X[26] + ( - (r2_6*X[26]) + (r1_6*X[16]))*dtr + - plom_fsqrtf((r2_6*X[26]))*d2_6+ plom_fsqrtf((r1_6*X[16]))*d1_6;
        s36 += ( + (r1_6*X[16]))*dtr + + plom_fsqrtf((r1_6*X[16]))*d1_6;
        X[6] = plom_clipf(f6);
        X[16] = plom_clipf(f16);
        const float r0_7 = plom_fdivf((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])), C7);
        const float r1_7 = P17;
        const float r2_7 = P27;
        const float r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        float za14, zb14;
        plom_normal_pair32(ns, &za14, &zb14);
        const float d0_7 = sqdt * za14;
        const float d1_7 = sqdt * zb14;
        float za15, zb15;
        plom_normal_pair32(ns, &za15, &zb15);
        const float d2_7 = sqdt * za15;
        const float d3_7 = sqdt * zb15;
        const float f7 = X[7] + ( - (r0_7*X[7]) + (r3_7))*dtr + - plom_fsqrtf((r0_7*X[7]))*d0_7+ plom_fsqrtf((r3_7))*d3_7;
        const float f17 = X[17] + ( - (r1_7*X[17]) + (r0_7*X[7]))*dtr + - plom_fsqrtf((r1_7*X[17]))*d1_7+ plom_fsqrtf((r0_7*X[7]))*d0_7;
        const float f27 = X[27] + ( - (r2_7*X[27]) + (r1_7*X[17]))*dtr + - plom_fsqrtf((r2_7*X[27]))*d2_7+ plom_fsqrtf((r1_7*X[17]))*d1_7;
        s37 += ( + (r1_7*X[17]))*dtr + + plom_fsqrtf((r1_7*X[17]))*d1_7;
        X[7] = plom_clipf(f7);
        X[17] = plom_clipf(f17);
        const float r0_8 = plom_fdivf((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])), C8);
        const float r1_8 = P18;
        const float r2_8 = P28;
        const float r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        float za16, zb16;
        plom_normal_pair32(ns, &za16, &zb16);
        const float d0_8 = sqdt * za16;
        const float d1_8 = sqdt * zb16;
        float za17, zb17;
        plom_normal_pair32(ns, &za17, &zb17);
        const float d2_8 = sqdt * za17;
        const float d3_8 = sqdt * zb17;
        const float f8 = X[8] + ( - (r0_8*X[8]) + (r3_8))*dtr + - plom_fsqrtf((r0_8*X[8]))*d0_8+ plom_fsqrtf((r3_8))*d3_8;
        const float f18 = X[18] + ( - (r1_8*X[18]) + (r0_8*X[8]))*dtr + - plom_fsqrtf((r1_8*X[18]))*d1_8+ plom_fsqrtf((r0_8*X[8]))*d0_8;
        const float f28 = X[28] + ( - (r2_8*X[28]) + (r1_8*X[18]))*dtr + - plom_fsqrtf((r2_8*X[28]))*d2_8+ plom_fsqrtf((r1_8*X[18]))*d1_8;
        s38 += ( + (r1_8*X[18]))*dtr + + plom_fsqrtf((r1_8*X[18]))*d1_8;
        X[8] = plom_clipf(f8);
        X[18] = plom_clipf(f18);
        const float r0_9 = plom_fdivf((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])), C9);
        const float r1_9 = P19;
        const float r2_9 = P29;
        const float r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        float za18, zb18;
        plom_normal_pair32(ns, &za18, &zb18);
        const float d0_9 = sqdt * za18;
        const float d1_9 = sqdt * zb18;
        float za19, zb19;
        plom_normal_pair32(ns, &za19, &zb19);
        const float d2_9 = sqdt * za19;
        const float d3_9 = sqdt * zb19;
        const float f9 = X[9] + ( - (r0_9*X[9]) + (r3_9))*dtr + - plom_fsqrtf((r0_9*X[9]))*d0_9+ plom_fsqrtf((r3_9))*d3_9;
        const float f19 = X[19] + ( - (r1_9*X[19]) + (r0_9*X[9]))*dtr + - plom_fsqrtf((r1_9*X[19]))*d1_9+ plom_fsqrtf((r0_9*X[9]))*d0_9;
        const float f29 = X[29] + ( - (r2_9*X[29]) + (r1_9*X[19]))*dtr + - plom_fsqrtf((r2_9*X[29]))*d2_9+ plom_fsqrtf((r1_9*X[19]))*d1_9;
        s39 += ( + (r1_9*X[19]))*dtr + + plom_fsqrtf((r1_9*X[19]))*d1_9;
        X[9] = plom_clipf(f9);
        X[19] = plom_clipf(f19);
        X[20] = plom_clipf(f20);
        X[21] = plom_clipf(f21);
        X[22] = plom_clipf(f22);
        X[23] = plom_clipf(f23);
        X[24] = plom_clipf(f24);
        X[25] = plom_clipf(f25);
        X[26] = plom_clipf(f26);
        X[27] = plom_clipf(f27);
        X[28] = plom_clipf(f28);
        X[29] = plom_clipf(f29);
        const float f30 = X[30] + s30;
        X[30] = plom_clipf(f30);
        const float f31 = X[31] + s31;
        X[31] = plom_clipf(f31);
        const float f32 = X[32] + s32;
        X[32] = plom_clipf(f32);
        const float f33 = X[33] + s33;
        X[33] = plom_clipf(f33);
        const float f34 = X[34] + s34;
        X[34] = plom_clipf(f34);
        const float f35 = X[35] + s35;
        X[35] = plom_clipf(f35);
        const float f36 = X[36] + s36;
        X[36] = plom_clipf(f36);
        const float f37 = X[37] + s37;
        X[37] = plom_clipf(f37);
        const float f38 = X[38] + s38;
        X[38] = plom_clipf(f38);
        const float f39 = X[39] + s39;
        X[39] = plom_clipf(f39);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_interval_no_dem_sto_no_env_sto_f32(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
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
    const float P20 = PLOM_UNIFORM32(pv[20]);
    const float P21 = PLOM_UNIFORM32(pv[21]);
    const float P22 = PLOM_UNIFORM32(pv[22]);
    const float P23 = PLOM_UNIFORM32(pv[23]);
    const float P24 = PLOM_UNIFORM32(pv[24]);
    const float P25 = PLOM_UNIFORM32(pv[25]);
    const float P26 = PLOM_UNIFORM32(pv[26]);
    const float P27 = PLOM_UNIFORM32(pv[27]);
    const float P28 = PLOM_UNIFORM32(pv[28]);
    const float P29 = PLOM_UNIFORM32(pv[29]);
    const float P30 = PLOM_UNIFORM32(pv[30]);
    const float P31 = PLOM_UNIFORM32(pv[31]);
    const float P32 = PLOM_UNIFORM32(pv[32]);
    const float P33 = PLOM_UNIFORM32(pv[33]);
    const float P34 = PLOM_UNIFORM32(pv[34]);
    const float P35 = PLOM_UNIFORM32(pv[35]);
    const float P36 = PLOM_UNIFORM32(pv[36]);
    const float P37 = PLOM_UNIFORM32(pv[37]);
    const float P38 = PLOM_UNIFORM32(pv[38]);
    const float P39 = PLOM_UNIFORM32(pv[39]);
    const float P80 = PLOM_UNIFORM32(pv[80]);
    const float P81 = PLOM_UNIFORM32(pv[81]);
    const float P82 = PLOM_UNIFORM32(pv[82]);
    const float P83 = PLOM_UNIFORM32(pv[83]);
    const float P84 = PLOM_UNIFORM32(pv[84]);
    const float P85 = PLOM_UNIFORM32(pv[85]);
    const float P86 = PLOM_UNIFORM32(pv[86]);
    const float P87 = PLOM_UNIFORM32(pv[87]);
    const float P88 = PLOM_UNIFORM32(pv[88]);
    const float P89 = PLOM_UNIFORM32(pv[89]);
    const float P90 = PLOM_UNIFORM32(pv[90]);
    const float P91 = PLOM_UNIFORM32(pv[91]);
    const float P92 = PLOM_UNIFORM32(pv[92]);
    const float P93 = PLOM_UNIFORM32(pv[93]);
    const float P94 = PLOM_UNIFORM32(pv[94]);
    const float P95 = PLOM_UNIFORM32(pv[95]);
    const float P96 = PLOM_UNIFORM32(pv[96]);
    const float P97 = PLOM_UNIFORM32(pv[97]);
    const float P98 = PLOM_UNIFORM32(pv[98]);
    const float P99 = PLOM_UNIFORM32(pv[99]);
    const float P100 = PLOM_UNIFORM32(pv[100]);
    const float P101 = PLOM_UNIFORM32(pv[101]);
    const float P102 = PLOM_UNIFORM32(pv[102]);
    const float P103 = PLOM_UNIFORM32(pv[103]);
    const float P104 = PLOM_UNIFORM32(pv[104]);
    const float P105 = PLOM_UNIFORM32(pv[105]);
    const float P106 = PLOM_UNIFORM32(pv[106]);
    const float P107 = PLOM_UNIFORM32(pv[107]);
    const float P108 = PLOM_UNIFORM32(pv[108]);
    const float P109 = PLOM_UNIFORM32(pv[109]);
    const float P110 = PLOM_UNIFORM32(pv[110]);
    const float P111 = PLOM_UNIFORM32(pv[111]);
    const float P112 = PLOM_UNIFORM32(pv[112]);
    const float P113 = PLOM_UNIFORM32(pv[113]);
    const float P114 = PLOM_UNIFORM32(pv[114]);
    const float P115 = PLOM_UNIFORM32(pv[115]);
    const float P116 = PLOM_UNIFORM32(pv[116]);
    const float P117 = PLOM_UNIFORM32(pv[117]);
    const float P118 = PLOM_UNIFORM32(pv[118]);
    const float P119 = PLOM_UNIFORM32(pv[119]);
    const float P120 = PLOM_UNIFORM32(pv[120]);
    const float P121 = PLOM_UNIFORM32(pv[121]);
    const float P122 = PLOM_UNIFORM32(pv[122]);
    const float P123 = PLOM_UNIFORM32(pv[123]);
    const float P124 = PLOM_UNIFORM32(pv[124]);
    const float P125 = PLOM_UNIFORM32(pv[125]);
    const float P126 = PLOM_UNIFORM32(pv[126]);
    const float P127 = PLOM_UNIFORM32(pv[127]);
    const float P128 = PLOM_UNIFORM32(pv[128]);
    const float P129 = PLOM_UNIFORM32(pv[129]);
    const float P130 = PLOM_UNIFORM32(pv[130]);
    const float P131 = PLOM_UNIFORM32(pv[131]);
    const float P132 = PLOM_UNIFORM32(pv[132]);
    const float P133 = PLOM_UNIFORM32(pv[133]);
    const float P134 = PLOM_UNIFORM32(pv[134]);
    const float P135 = PLOM_UNIFORM32(pv[135]);
    const float P136 = PLOM_UNIFORM32(pv[136]);
    const float P137 = PLOM_UNIFORM32(pv[137]);
    const float P138 = PLOM_UNIFORM32(pv[138]);
    const float P139 = PLOM_UNIFORM32(pv[139]);
    const float P140 = PLOM_UNIFORM32(pv[140]);
    const float P141 = PLOM_UNIFORM32(pv[141]);
    const float P142 = PLOM_UNIFORM32(pv[142]);
    const float P143 = PLOM_UNIFORM32(pv[143]);
    const float P144 = PLOM_UNIFORM32(pv[144]);
    const float P145 = PLOM_UNIFORM32(pv[145]);
    const float P146 = PLOM_UNIFORM32(pv[146]);
    const float P147 = PLOM_UNIFORM32(pv[147]);
    const float P148 = PLOM_UNIFORM32(pv[148]);
    const float P149 = PLOM_UNIFORM32(pv[149]);
    const float P150 = PLOM_UNIFORM32(pv[150]);
    const float P151 = PLOM_UNIFORM32(pv[151]);
    const float P152 = PLOM_UNIFORM32(pv[152]);
    const float P153 = PLOM_UNIFORM32(pv[153]);
    const float P154 = PLOM_UNIFORM32(pv[154]);
    const float P155 = PLOM_UNIFORM32(pv[155]);
    const float P156 = PLOM_UNIFORM32(pv[156]);
    const float P157 = PLOM_UNIFORM32(pv[157]);
    const float P158 = PLOM_UNIFORM32(pv[158]);
    const float P159 = PLOM_UNIFORM32(pv[159]);
    const float P160 = PLOM_UNIFORM32(pv[160]);
    const float P161 = PLOM_UNIFORM32(pv[161]);
    const float P162 = PLOM_UNIFORM32(pv[162]);
    const float P163 = PLOM_UNIFORM32(pv[163]);
    const float P164 = PLOM_UNIFORM32(pv[164]);
    const float P165 = PLOM_UNIFORM32(pv[165]);
    const float P166 = PLOM_UNIFORM32(pv[166]);
    const float P167 = PLOM_UNIFORM32(pv[167]);
    const float P168 = PLOM_UNIFORM32(pv[168]);
    const float P169 = PLOM_UNIFORM32(pv[169]);
    const float P170 = PLOM_UNIFORM32(pv[170]);
    const float P171 = PLOM_UNIFORM32(pv[171]);
    const float P172 = PLOM_UNIFORM32(pv[172]);
    const float P173 = PLOM_UNIFORM32(pv[173]);
    const float P174 = PLOM_UNIFORM32(pv[174]);
    const float P175 = PLOM_UNIFORM32(pv[175]);
    const float P176 = PLOM_UNIFORM32(pv[176]);
    const float P177 = PLOM_UNIFORM32(pv[177]);
    const float P178 = PLOM_UNIFORM32(pv[178]);
    const float P179 = PLOM_UNIFORM32(pv[179]);
    X[30] = 0.0f; /* reset_inc */
    X[31] = 0.0f; /* reset_inc */
    X[32] = 0.0f; /* reset_inc */
    X[33] = 0.0f; /* reset_inc */
    X[34] = 0.0f; /* reset_inc */
    X[35] = 0.0f; /* reset_inc */
    X[36] = 0.0f; /* reset_inc */
    X[37] = 0.0f; /* reset_inc */
    X[38] = 0.0f; /* reset_inc */
    X[39] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        const float C1 = (float)plom_cov_ld(cvk + 1);
        const float C2 = (float)plom_cov_ld(cvk + 2);
        const float C3 = (float)plom_cov_ld(cvk + 3);
        const float C4 = (float)plom_cov_ld(cvk + 4);
        const float C5 = (float)plom_cov_ld(cvk + 5);
        const float C6 = (float)plom_cov_ld(cvk + 6);
        const float C7 = (float)plom_cov_ld(cvk + 7);
        const float C8 = (float)plom_cov_ld(cvk + 8);
        const float C9 = (float)plom_cov_ld(cvk + 9);
        float s30 = 0.0f;
        float s31 = 0.0f;
        float s32 = 0.0f;
        float s33 = 0.0f;
        float s34 = 0.0f;
        float s35 = 0.0f;
        float s36 = 0.0f;
        float s37 = 0.0f;
        float s38 = 0.0f;
        float s39 = 0.0f;
        const float r0_0 = plom_fdivf((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])), C0);
        const float r1_0 = P10;
        const float r2_0 = P20;
        const float r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r3_0))*dtr;
        const float f10 = X[10] + ( - (r1_0*X[10]) + (r0_0*X[0]))*dtr;
        const float f20 = X[20] + ( - (r2_0*X[20]) + (r1_0*X[10]))*dtr;
        s30 += ( + (r1_0*X[10]))*dtr;
        X[0] = plom_clipf(f0);
        X[10] = plom_clipf(f10);
        const float r0_1 = plom_fdivf((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])), C1);
        const float r1_1 = P11;
        const float r2_1 = P21;
        const float r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r3_1))*dtr;
        const float f11 = X[11] + ( - (r1_1*X[11]) + (r0_1*X[1]))*dtr;
        const float f21 = X[21] + ( - (r2_1*X[21]) + (r1_1*X[11]))*dtr;
        s31 += ( + (r1_1*X[11]))*dtr;
        X[1] = plom_clipf(f1);
        X[11] = plom_clipf(f11);
        const float r0_2 = plom_fdivf((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])), C2);
        const float r1_2 = P12;
        const float r2_2 = P22;
        const float r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r3_2))*dtr;
        const float f12 = X[12] + ( - (r1_2*X[12]) + (r0_2*X[2]))*dtr;
        const float f22 = X[22] + ( - (r2_2*X[22]) + (r1_2*X[12]))*dtr;
        s32 += ( + (r1_2*X[12]))*dtr;
        X[2] = plom_clipf(f2);
        X[12] = plom_clipf(f12);
        const float r0_3 = plom_fdivf((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])), C3);
        const float r1_3 = P13;
        const float r2_3 = P23;
        const float r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r3_3))*dtr;
        const float f13 = X[13] + ( - (r1_3*X[13]) + (r0_3*X[3]))*dtr;
        const float f23 = X[23] + ( - (r2_3*X[23]) + (r1_3*X[13]))*dtr;
        s33 += ( + (r1_3*X[13]))*dtr;
        X[3] = plom_clipf(f3);
        X[13] = plom_clipf(f13);
        const float r0_4 = plom_fdivf((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])), C4);
        const float r1_4 = P14;
        const float r2_4 = P24;
        const float r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        const float f4 = X[4] + ( - (r0_4*X[4]) + (r3_4))*dtr;
        const float f14 = X[14] + ( - (r1_4*X[14]) + (r0_4*X[4]))*dtr;
        const float f24 = X[24] + ( - (r2_4*X[24]) + (r1_4*X[14]))*dtr;
        s34 += ( + (r1_4*X[14]))*dtr;
        X[4] = plom_clipf(f4);
        X[14] = plom_clipf(f14);
        const float r0_5 = plom_fdivf((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])), C5);
        const float r1_5 = P15;
        const float r2_5 = P25;
        const float r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        const float f5 = X[5] + ( - (r0_5*X[5]) + (r3_5))*dtr;
        const float f15 = X[15] + ( - (r1_5*X[15]) + (r0_5*X[5]))*dtr;
        const float f25 = X[25] + ( - (r2_5*X[25]) + (r1_5*X[15]))*dtr;
        s35 += ( + (r1_5*X[15]))*dtr;
        X[5] = plom_clipf(f5);
        X[15] = plom_clipf(f15);
        const float r0_6 = plom_fdivf((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])), C6);
        const float r1_6 = P16;
        const float r2_6 = P26;
        const float r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        const float f6 = X[6] + ( - (r0_6*X[6]) + (r3_6))*dtr;
        const float f16 = X[16] + ( - (r1_6*X[16]) + (r0_6*X[6]))*dtr;
        const float f26 = X[26] + ( - (r2_6*X[26]) + (r1_6*X[16]))*dtr;
        s36 += ( + (r1_6*X[16]))*dtr;
        X[6] = plom_clipf(f6);
        X[16] = plom_clipf(f16);
        const float r0_7 = plom_fdivf((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])), C7);
        const float r1_7 = P17;
        const float r2_7 = P27;
        const float r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        const float f7 = X[7] + ( - (r0_7*X[7]) + (r3_7))*dtr;
        const float f17 = X[17] + ( - (r1_7*X[17]) + (r0_7*X[7]))*dtr;
        const float f27 = X[27] + ( - (r2_7*X[27]) + (r1_7*X[17]))*dtr;
        s37 += ( + (r1_7*X[17]))*dtr;
        X[7] = plom_clipf(f7);
        X[17] = plom_clipf(f17);
        const float r0_8 = plom_fdivf((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])), C8);
        const float r1_8 = P18;
        const float r2_8 = P28;
        const float r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        const float f8 = X[8] + ( - (r0_8*X[8]) + (r3_8))*dtr;
        const float f18 = X[18] + ( - (r1_8*X[18]) + (r0_8*X[8]))*dtr;
        const float f28 = X[28] + ( - (r2_8*X[28]) + (r1_8*X[18]))*dtr;
        s38 += ( + (r1_8*X[18]))*dtr;
        X[8] = plom_clipf(f8);
        X[18] = plom_clipf(f18);
        const float r0_9 = plom_fdivf((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])), C9);
        const float r1_9 = P19;
        const float r2_9 = P29;
        const float r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        const float f9 = X[9] + ( - (r0_9*X[9]) + (r3_9))*dtr;
        const float f19 = X[19] + ( - (r1_9*X[19]) + (r0_9*X[9]))*dtr;
        const float f29 = X[29] + ( - (r2_9*X[29]) + (r1_9*X[19]))*dtr;
        s39 += ( + (r1_9*X[19]))*dtr;
        X[9] = plom_clipf(f9);
        X[19] = plom_clipf(f19);
        X[20] = plom_clipf(f20);
        X[21] = plom_clipf(f21);
        X[22] = plom_clipf(f22);
        X[23] = plom_clipf(f23);
        X[24] = plom_clipf(f24);
        X[25] = plom_clipf(f25);
        X[26] = plom_clipf(f26);
        X[27] = plom_clipf(f27);
        X[28] = plom_clipf(f28);
        X[29] = plom_clipf(f29);
        const float f30 = X[30] + s30;
        X[30] = plom_clipf(f30);
        const float f31 = X[31] + s31;
        X[31] = plom_clipf(f31);
        const float f32 = X[32] + s32;
        X[32] = plom_clipf(f32);
        const float f33 = X[33] + s33;
        X[33] = plom_clipf(f33);
        const float f34 = X[34] + s34;
        X[34] = plom_clipf(f34);
        const float f35 = X[35] + s35;
        X[35] = plom_clipf(f35);
        const float f36 = X[36] + s36;
        X[36] = plom_clipf(f36);
        const float f37 = X[37] + s37;
        X[37] = plom_clipf(f37);
        const float f38 = X[38] + s38;
        X[38] = plom_clipf(f38);
        const float f39 = X[39] + s39;
        X[39] = plom_clipf(f39);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_proj2obs_f32(const float *X, const double *pvd, const double *cvd, double *obs)
{
    (void)pvd; (void)cvd; (void)X;
    obs[0] = (double)X[30];
    obs[1] = (double)X[31];
    obs[2] = (double)X[32];
    obs[3] = (double)X[33];
    obs[4] = (double)X[34];
    obs[5] = (double)X[35];
    obs[6] = (double)X[36];
    obs[7] = (double)X[37];
    obs[8] = (double)X[38];
    obs[9] = (double)X[39];
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_interval_psr_full_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
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
    const double P20 = lp[20];
    const double P21 = lp[21];
    const double P22 = lp[22];
    const double P23 = lp[23];
    const double P24 = lp[24];
    const double P25 = lp[25];
    const double P26 = lp[26];
    const double P27 = lp[27];
    const double P28 = lp[28];
    const double P29 = lp[29];
    const double P30 = lp[30];
    const double P31 = lp[31];
    const double P32 = lp[32];
    const double P33 = lp[33];
    const double P34 = lp[34];
    const double P35 = lp[35];
    const double P36 = lp[36];
    const double P37 = lp[37];
    const double P38 = lp[38];
    const double P39 = lp[39];
    const double P40 = lp[40];
    const double P41 = lp[41];
    const double P42 = lp[42];
    const double P43 = lp[43];
    const double P44 = lp[44];
    const double P45 = lp[45];
    const double P46 = lp[46];
    const double P47 = lp[47];
    const double P48 = lp[48];
    const double P49 = lp[49];
    const double P80 = PLOM_UNIFORM64(pv[80]);
    const double P81 = PLOM_UNIFORM64(pv[81]);
    const double P82 = PLOM_UNIFORM64(pv[82]);
    const double P83 = PLOM_UNIFORM64(pv[83]);
    const double P84 = PLOM_UNIFORM64(pv[84]);
    const double P85 = PLOM_UNIFORM64(pv[85]);
    const double P86 = PLOM_UNIFORM64(pv[86]);
    const double P87 = PLOM_UNIFORM64(pv[87]);
    const double P88 = PLOM_UNIFORM64(pv[88]);
    const double P89 = PLOM_UNIFORM64(pv[89]);
    const double P90 = PLOM_UNIFORM64(pv[90]);
    const double P91 = PLOM_UNIFORM64(pv[91]);
    const double P92 = PLOM_UNIFORM64(pv[92]);
    const double P93 = PLOM_UNIFORM64(pv[93]);
    const double P94 = PLOM_UNIFORM64(pv[94]);
    const double P95 = PLOM_UNIFORM64(pv[95]);
    const double P96 = PLOM_UNIFORM64(pv[96]);
    const double P97 = PLOM_UNIFORM64(pv[97]);
    const double P98 = PLOM_UNIFORM64(pv[98]);
    const double P99 = PLOM_UNIFORM64(pv[99]);
    const double P100 = PLOM_UNIFORM64(pv[100]);
    const double P101 = PLOM_UNIFORM64(pv[101]);
    const double P102 = PLOM_UNIFORM64(pv[102]);
    const double P103 = PLOM_UNIFORM64(pv[103]);
    const double P104 = PLOM_UNIFORM64(pv[104]);
    const double P105 = PLOM_UNIFORM64(pv[105]);
    const double P106 = PLOM_UNIFORM64(pv[106]);
    const double P107 = PLOM_UNIFORM64(pv[107]);
    const double P108 = PLOM_UNIFORM64(pv[108]);
    const double P109 = PLOM_UNIFORM64(pv[109]);
    const double P110 = PLOM_UNIFORM64(pv[110]);
    const double P111 = PLOM_UNIFORM64(pv[111]);
    const double P112 = PLOM_UNIFORM64(pv[112]);
    const double P113 = PLOM_UNIFORM64(pv[113]);
    const double P114 = PLOM_UNIFORM64(pv[114]);
    const double P115 = PLOM_UNIFORM64(pv[115]);
    const double P116 = PLOM_UNIFORM64(pv[116]);
    const double P117 = PLOM_UNIFORM64(pv[117]);
    const double P118 = PLOM_UNIFORM64(pv[118]);
    const double P119 = PLOM_UNIFORM64(pv[119]);
    const double P120 = PLOM_UNIFORM64(pv[120]);
    const double P121 = PLOM_UNIFORM64(pv[121]);
    const double P122 = PLOM_UNIFORM64(pv[122]);
    const double P123 = PLOM_UNIFORM64(pv[123]);
    const double P124 = PLOM_UNIFORM64(pv[124]);
    const double P125 = PLOM_UNIFORM64(pv[125]);
    const double P126 = PLOM_UNIFORM64(pv[126]);
    const double P127 = PLOM_UNIFORM64(pv[127]);
    const double P128 = PLOM_UNIFORM64(pv[128]);
    const double P129 = PLOM_UNIFORM64(pv[129]);
    const double P130 = PLOM_UNIFORM64(pv[130]);
    const double P131 = PLOM_UNIFORM64(pv[131]);
    const double P132 = PLOM_UNIFORM64(pv[132]);
    const double P133 = PLOM_UNIFORM64(pv[133]);
    const double P134 = PLOM_UNIFORM64(pv[134]);
    const double P135 = PLOM_UNIFORM64(pv[135]);
    const double P136 = PLOM_UNIFORM64(pv[136]);
    const double P137 = PLOM_UNIFORM64(pv[137]);
    const double P138 = PLOM_UNIFORM64(pv[138]);
    const double P139 = PLOM_UNIFORM64(pv[139]);
    const double P140 = PLOM_UNIFORM64(pv[140]);
    const double P141 = PLOM_UNIFORM64(pv[141]);
    const double P142 = PLOM_UNIFORM64(pv[142]);
    const double P143 = PLOM_UNIFORM64(pv[143]);
    const double P144 = PLOM_UNIFORM64(pv[144]);
    const double P145 = PLOM_UNIFORM64(pv[145]);
    const double P146 = PLOM_UNIFORM64(pv[146]);
    const double P147 = PLOM_UNIFORM64(pv[147]);
    const double P148 = PLOM_UNIFORM64(pv[148]);
    const double P149 = PLOM_UNIFORM64(pv[149]);
    const double P150 = PLOM_UNIFORM64(pv[150]);
    const double P151 = PLOM_UNIFORM64(pv[151]);
    const double P152 = PLOM_UNIFORM64(pv[152]);
    const double P153 = PLOM_UNIFORM64(pv[153]);
    const double P154 = PLOM_UNIFORM64(pv[154]);
    const double P155 = PLOM_UNIFORM64(pv[155]);
    const double P156 = PLOM_UNIFORM64(pv[156]);
    const double P157 = PLOM_UNIFORM64(pv[157]);
    const double P158 = PLOM_UNIFORM64(pv[158]);
    const double P159 = PLOM_UNIFORM64(pv[159]);
    const double P160 = PLOM_UNIFORM64(pv[160]);
    const double P161 = PLOM_UNIFORM64(pv[161]);
    const double P162 = PLOM_UNIFORM64(pv[162]);
    const double P163 = PLOM_UNIFORM64(pv[163]);
    const double P164 = PLOM_UNIFORM64(pv[164]);
    const double P165 = PLOM_UNIFORM64(pv[165]);
    const double P166 = PLOM_UNIFORM64(pv[166]);
    const double P167 = PLOM_UNIFORM64(pv[167]);
    const double P168 = PLOM_UNIFORM64(pv[168]);
    const double P169 = PLOM_UNIFORM64(pv[169]);
    const double P170 = PLOM_UNIFORM64(pv[170]);
    const double P171 = PLOM_UNIFORM64(pv[171]);
    const double P172 = PLOM_UNIFORM64(pv[172]);
    const double P173 = PLOM_UNIFORM64(pv[173]);
    const double P174 = PLOM_UNIFORM64(pv[174]);
    const double P175 = PLOM_UNIFORM64(pv[175]);
    const double P176 = PLOM_UNIFORM64(pv[176]);
    const double P177 = PLOM_UNIFORM64(pv[177]);
    const double P178 = PLOM_UNIFORM64(pv[178]);
    const double P179 = PLOM_UNIFORM64(pv[179]);
    X[30] = 0.0; /* reset_inc */
    X[31] = 0.0; /* reset_inc */
    X[32] = 0.0; /* reset_inc */
    X[33] = 0.0; /* reset_inc */
    X[34] = 0.0; /* reset_inc */
    X[35] = 0.0; /* reset_inc */
    X[36] = 0.0; /* reset_inc */
    X[37] = 0.0; /* reset_inc */
    X[38] = 0.0; /* reset_inc */
    X[39] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double C4 = plom_cov_ld(cvk + 4);
        const double C5 = plom_cov_ld(cvk + 5);
        const double C6 = plom_cov_ld(cvk + 6);
        const double C7 = plom_cov_ld(cvk + 7);
        const double C8 = plom_cov_ld(cvk + 8);
        const double C9 = plom_cov_ld(cvk + 9);
        double n0_0 = 0.0;
        double n1_0 = 0.0;
        double n2_0 = 0.0;
        double n3_0 = 0.0;
        double n0_1 = 0.0;
        double n1_1 = 0.0;
        double n2_1 = 0.0;
        double n3_1 = 0.0;
        double n0_2 = 0.0;
        double n1_2 = 0.0;
        double n2_2 = 0.0;
        double n3_2 = 0.0;
        double n0_3 = 0.0;
        double n1_3 = 0.0;
        double n2_3 = 0.0;
        double n3_3 = 0.0;
        double n0_4 = 0.0;
        double n1_4 = 0.0;
        double n2_4 = 0.0;
        double n3_4 = 0.0;
        double n0_5 = 0.0;
        double n1_5 = 0.0;
        double n2_5 = 0.0;
        double n3_5 = 0.0;
        double n0_6 = 0.0;
        double n1_6 = 0.0;
        double n2_6 = 0.0;
        double n3_6 = 0.0;
        double n0_7 = 0.0;
        double n1_7 = 0.0;
        double n2_7 = 0.0;
        double n3_7 = 0.0;
        double n0_8 = 0.0;
        double n1_8 = 0.0;
        double n2_8 = 0.0;
        double n3_8 = 0.0;
        double n0_9 = 0.0;
        double n1_9 = 0.0;
        double n2_9 = 0.0;
        double n3_9 = 0.0;
        double stay0 = 0.0;
        double stay1 = 0.0;
        double stay2 = 0.0;
        double stay3 = 0.0;
        double stay4 = 0.0;
        double stay5 = 0.0;
        double stay6 = 0.0;
        double stay7 = 0.0;
        double stay8 = 0.0;
        double stay9 = 0.0;
        double stay10 = 0.0;
        double stay11 = 0.0;
        double stay12 = 0.0;
        double stay13 = 0.0;
        double stay14 = 0.0;
        double stay15 = 0.0;
        double stay16 = 0.0;
        double stay17 = 0.0;
        double stay18 = 0.0;
        double stay19 = 0.0;
        double stay20 = 0.0;
        double stay21 = 0.0;
        double stay22 = 0.0;
        double stay23 = 0.0;
        double stay24 = 0.0;
        double stay25 = 0.0;
        double stay26 = 0.0;
        double stay27 = 0.0;
        double stay28 = 0.0;
        double stay29 = 0.0;
        (void)n0_0; (void)n1_0; (void)n2_0; (void)n3_0; (void)n0_1; (void)n1_1; (void)n2_1; (void)n3_1; (void)n0_2; (void)n1_2; (void)n2_2; (void)n3_2; (void)n0_3; (void)n1_3; (void)n2_3; (void)n3_3; (void)n0_4; (void)n1_4; (void)n2_4; (void)n3_4; (void)n0_5; (void)n1_5; (void)n2_5; (void)n3_5; (void)n0_6; (void)n1_6; (void)n2_6; (void)n3_6; (void)n0_7; (void)n1_7; (void)n2_7; (void)n3_7; (void)n0_8; (void)n1_8; (void)n2_8; (void)n3_8; (void)n0_9; (void)n1_9; (void)n2_9; (void)n3_9;
        const double r0_0 = ((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])) / C0);
        const double r1_0 = P10;
        const double r2_0 = P20;
        const double r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        const double r0_1 = ((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])) / C1);
        const double r1_1 = P11;
        const double r2_1 = P21;
        const double r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        const double r0_2 = ((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])) / C2);
        const double r1_2 = P12;
        const double r2_2 = P22;
        const double r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        const double r0_3 = ((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])) / C3);
        const double r1_3 = P13;
        const double r2_3 = P23;
        const double r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        const double r0_4 = ((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])) / C4);
        const double r1_4 = P14;
        const double r2_4 = P24;
        const double r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        const double r0_5 = ((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])) / C5);
        const double r1_5 = P15;
        const double r2_5 = P25;
        const double r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        const double r0_6 = ((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])) / C6);
        const double r1_6 = P16;
        const double r2_6 = P26;
        const double r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        const double r0_7 = ((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])) / C7);
        const double r1_7 = P17;
        const double r2_7 = P27;
        const double r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        const double r0_8 = ((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])) / C8);
        const double r1_8 = P18;
        const double r2_8 = P28;
        const double r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        const double r0_9 = ((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])) / C9);
        const double r1_9 = P19;
        const double r2_9 = P29;
        const double r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        const double sd0_0 = P40;
        const double v0_0 = sd0_0 * sd0_0;
        const double xi0_0 = (v0_0 * plom_gamma_draw(ns, dt / v0_0)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_0 * xi0_0) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_0 * xi0_0) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[0]), p, cnt);
            n0_0 = cnt[0];
            stay0 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_0 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_0 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[10]), p, cnt);
            n1_0 = cnt[0];
            stay10 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_0 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_0 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[20]), p, cnt);
            n2_0 = cnt[0];
            stay20 = cnt[1];
        }
        n3_0 = plom_poisson_draw(ns, r3_0 * dt);
        const double sd0_1 = P41;
        const double v0_1 = sd0_1 * sd0_1;
        const double xi0_1 = (v0_1 * plom_gamma_draw(ns, dt / v0_1)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_1 * xi0_1) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_1 * xi0_1) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[1]), p, cnt);
            n0_1 = cnt[0];
            stay1 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_1 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_1 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[11]), p, cnt);
            n1_1 = cnt[0];
            stay11 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_1 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_1 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[21]), p, cnt);
            n2_1 = cnt[0];
            stay21 = cnt[1];
        }
        n3_1 = plom_poisson_draw(ns, r3_1 * dt);
        const double sd0_2 = P42;
        const double v0_2 = sd0_2 * sd0_2;
        const double xi0_2 = (v0_2 * plom_gamma_draw(ns, dt / v0_2)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_2 * xi0_2) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_2 * xi0_2) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[2]), p, cnt);
            n0_2 = cnt[0];
            stay2 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_2 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_2 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[12]), p, cnt);
            n1_2 = cnt[0];
            stay12 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_2 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_2 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[22]), p, cnt);
            n2_2 = cnt[0];
            stay22 = cnt[1];
        }
        n3_2 = plom_poisson_draw(ns, r3_2 * dt);
        const double sd0_3 = P43;
        const double v0_3 = sd0_3 * sd0_3;
        const double xi0_3 = (v0_3 * plom_gamma_draw(ns, dt / v0_3)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_3 * xi0_3) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_3 * xi0_3) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[3]), p, cnt);
            n0_3 = cnt[0];
            stay3 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_3 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_3 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[13]), p, cnt);
            n1_3 = cnt[0];
            stay13 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_3 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_3 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[23]), p, cnt);
            n2_3 = cnt[0];
            stay23 = cnt[1];
        }
        n3_3 = plom_poisson_draw(ns, r3_3 * dt);
        const double sd0_4 = P44;
        const double v0_4 = sd0_4 * sd0_4;
        const double xi0_4 = (v0_4 * plom_gamma_draw(ns, dt / v0_4)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_4 * xi0_4) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_4 * xi0_4) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[4]), p, cnt);
            n0_4 = cnt[0];
            stay4 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_4 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_4 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[14]), p, cnt);
            n1_4 = cnt[0];
            stay14 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_4 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_4 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[24]), p, cnt);
            n2_4 = cnt[0];
            stay24 = cnt[1];
        }
        n3_4 = plom_poisson_draw(ns, r3_4 * dt);
        const double sd0_5 = P45;
        const double v0_5 = sd0_5 * sd0_5;
        const double xi0_5 = (v0_5 * plom_gamma_draw(ns, dt / v0_5)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_5 * xi0_5) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_5 * xi0_5) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[5]), p, cnt);
            n0_5 = cnt[0];
            stay5 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_5 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_5 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[15]), p, cnt);
            n1_5 = cnt[0];
            stay15 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_5 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_5 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[25]), p, cnt);
            n2_5 = cnt[0];
            stay25 = cnt[1];
        }
        n3_5 = plom_poisson_draw(ns, r3_5 * dt);
        const double sd0_6 = P46;
        const double v0_6 = sd0_6 * sd0_6;
        const double xi0_6 = (v0_6 * plom_gamma_draw(ns, dt / v0_6)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_6 * xi0_6) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_6 * xi0_6) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[6]), p, cnt);
            n0_6 = cnt[0];
            stay6 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_6 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_6 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[16]), p, cnt);
            n1_6 = cnt[0];
            stay16 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_6 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_6 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[26]), p, cnt);
            n2_6 = cnt[0];
            stay26 = cnt[1];
        }
        n3_6 = plom_poisson_draw(ns, r3_6 * dt);
        const double sd0_7 = P47;
        const double v0_7 = sd0_7 * sd0_7;
        const double xi0_7 = (v0_7 * plom_gamma_draw(ns, dt / v0_7)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_7 * xi0_7) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_7 * xi0_7) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[7]), p, cnt);
            n0_7 = cnt[0];
            stay7 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_7 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_7 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[17]), p, cnt);
            n1_7 = cnt[0];
            stay17 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_7 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_7 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[27]), p, cnt);
            n2_7 = cnt[0];
            stay27 = cnt[1];
        }
        n3_7 = plom_poisson_draw(ns, r3_7 * dt);
        const double sd0_8 = P48;
        const double v0_8 = sd0_8 * sd0_8;
        const double xi0_8 = (v0_8 * plom_gamma_draw(ns, dt / v0_8)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_8 * xi0_8) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_8 * xi0_8) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[8]), p, cnt);
            n0_8 = cnt[0];
            stay8 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_8 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_8 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[18]), p, cnt);
            n1_8 = cnt[0];
            stay18 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_8 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_8 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[28]), p, cnt);
            n2_8 = cnt[0];
            stay28 = cnt[1];
        }
        n3_8 = plom_poisson_draw(ns, r3_8 * dt);
        const double sd0_9 = P49;
        const double v0_9 = sd0_9 * sd0_9;
        const double xi0_9 = (v0_9 * plom_gamma_draw(ns, dt / v0_9)) / dt;
        {
            double p[2], cnt[2];
            const double sum = (r0_9 * xi0_9) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_9 * xi0_9) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[9]), p, cnt);
            n0_9 = cnt[0];
            stay9 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_9 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_9 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[19]), p, cnt);
            n1_9 = cnt[0];
            stay19 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_9 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_9 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[29]), p, cnt);
            n2_9 = cnt[0];
            stay29 = cnt[1];
        }
        n3_9 = plom_poisson_draw(ns, r3_9 * dt);
        X[0] = stay0 + n3_0;
        X[10] = stay10 + n0_0;
        X[20] = stay20 + n1_0;
        X[1] = stay1 + n3_1;
        X[11] = stay11 + n0_1;
        X[21] = stay21 + n1_1;
        X[2] = stay2 + n3_2;
        X[12] = stay12 + n0_2;
        X[22] = stay22 + n1_2;
        X[3] = stay3 + n3_3;
        X[13] = stay13 + n0_3;
        X[23] = stay23 + n1_3;
        X[4] = stay4 + n3_4;
        X[14] = stay14 + n0_4;
        X[24] = stay24 + n1_4;
        X[5] = stay5 + n3_5;
        X[15] = stay15 + n0_5;
        X[25] = stay25 + n1_5;
        X[6] = stay6 + n3_6;
        X[16] = stay16 + n0_6;
        X[26] = stay26 + n1_6;
        X[7] = stay7 + n3_7;
        X[17] = stay17 + n0_7;
        X[27] = stay27 + n1_7;
        X[8] = stay8 + n3_8;
        X[18] = stay18 + n0_8;
        X[28] = stay28 + n1_8;
        X[9] = stay9 + n3_9;
        X[19] = stay19 + n0_9;
        X[29] = stay29 + n1_9;
        X[30] = X[30] + (n1_0);
        X[31] = X[31] + (n1_1);
        X[32] = X[32] + (n1_2);
        X[33] = X[33] + (n1_3);
        X[34] = X[34] + (n1_4);
        X[35] = X[35] + (n1_5);
        X[36] = X[36] + (n1_6);
        X[37] = X[37] + (n1_7);
        X[38] = X[38] + (n1_8);
        X[39] = X[39] + (n1_9);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_interval_psr_no_env_sto_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
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
    const double P20 = lp[20];
    const double P21 = lp[21];
    const double P22 = lp[22];
    const double P23 = lp[23];
    const double P24 = lp[24];
    const double P25 = lp[25];
    const double P26 = lp[26];
    const double P27 = lp[27];
    const double P28 = lp[28];
    const double P29 = lp[29];
    const double P30 = lp[30];
    const double P31 = lp[31];
    const double P32 = lp[32];
    const double P33 = lp[33];
    const double P34 = lp[34];
    const double P35 = lp[35];
    const double P36 = lp[36];
    const double P37 = lp[37];
    const double P38 = lp[38];
    const double P39 = lp[39];
    const double P80 = PLOM_UNIFORM64(pv[80]);
    const double P81 = PLOM_UNIFORM64(pv[81]);
    const double P82 = PLOM_UNIFORM64(pv[82]);
    const double P83 = PLOM_UNIFORM64(pv[83]);
    const double P84 = PLOM_UNIFORM64(pv[84]);
    const double P85 = PLOM_UNIFORM64(pv[85]);
    const double P86 = PLOM_UNIFORM64(pv[86]);
    const double P87 = PLOM_UNIFORM64(pv[87]);
    const double P88 = PLOM_UNIFORM64(pv[88]);
    const double P89 = PLOM_UNIFORM64(pv[89]);
    const double P90 = PLOM_UNIFORM64(pv[90]);
    const double P91 = PLOM_UNIFORM64(pv[91]);
    const double P92 = PLOM_UNIFORM64(pv[92]);
    const double P93 = PLOM_UNIFORM64(pv[93]);
    const double P94 = PLOM_UNIFORM64(pv[94]);
    const double P95 = PLOM_UNIFORM64(pv[95]);
    const double P96 = PLOM_UNIFORM64(pv[96]);
    const double P97 = PLOM_UNIFORM64(pv[97]);
    const double P98 = PLOM_UNIFORM64(pv[98]);
    const double P99 = PLOM_UNIFORM64(pv[99]);
    const double P100 = PLOM_UNIFORM64(pv[100]);
    const double P101 = PLOM_UNIFORM64(pv[101]);
    const double P102 = PLOM_UNIFORM64(pv[102]);
    const double P103 = PLOM_UNIFORM64(pv[103]);
    const double P104 = PLOM_UNIFORM64(pv[104]);
    const double P105 = PLOM_UNIFORM64(pv[105]);
    const double P106 = PLOM_UNIFORM64(pv[106]);
    const double P107 = PLOM_UNIFORM64(pv[107]);
    const double P108 = PLOM_UNIFORM64(pv[108]);
    const double P109 = PLOM_UNIFORM64(pv[109]);
    const double P110 = PLOM_UNIFORM64(pv[110]);
    const double P111 = PLOM_UNIFORM64(pv[111]);
    const double P112 = PLOM_UNIFORM64(pv[112]);
    const double P113 = PLOM_UNIFORM64(pv[113]);
    const double P114 = PLOM_UNIFORM64(pv[114]);
    const double P115 = PLOM_UNIFORM64(pv[115]);
    const double P116 = PLOM_UNIFORM64(pv[116]);
    const double P117 = PLOM_UNIFORM64(pv[117]);
    const double P118 = PLOM_UNIFORM64(pv[118]);
    const double P119 = PLOM_UNIFORM64(pv[119]);
    const double P120 = PLOM_UNIFORM64(pv[120]);
    const double P121 = PLOM_UNIFORM64(pv[121]);
    const double P122 = PLOM_UNIFORM64(pv[122]);
    const double P123 = PLOM_UNIFORM64(pv[123]);
    const double P124 = PLOM_UNIFORM64(pv[124]);
    const double P125 = PLOM_UNIFORM64(pv[125]);
    const double P126 = PLOM_UNIFORM64(pv[126]);
    const double P127 = PLOM_UNIFORM64(pv[127]);
    const double P128 = PLOM_UNIFORM64(pv[128]);
    const double P129 = PLOM_UNIFORM64(pv[129]);
    const double P130 = PLOM_UNIFORM64(pv[130]);
    const double P131 = PLOM_UNIFORM64(pv[131]);
    const double P132 = PLOM_UNIFORM64(pv[132]);
    const double P133 = PLOM_UNIFORM64(pv[133]);
    const double P134 = PLOM_UNIFORM64(pv[134]);
    const double P135 = PLOM_UNIFORM64(pv[135]);
    const double P136 = PLOM_UNIFORM64(pv[136]);
    const double P137 = PLOM_UNIFORM64(pv[137]);
    const double P138 = PLOM_UNIFORM64(pv[138]);
    const double P139 = PLOM_UNIFORM64(pv[139]);
    const double P140 = PLOM_UNIFORM64(pv[140]);
    const double P141 = PLOM_UNIFORM64(pv[141]);
    const double P142 = PLOM_UNIFORM64(pv[142]);
    const double P143 = PLOM_UNIFORM64(pv[143]);
    const double P144 = PLOM_UNIFORM64(pv[144]);
    const double P145 = PLOM_UNIFORM64(pv[145]);
    const double P146 = PLOM_UNIFORM64(pv[146]);
    const double P147 = PLOM_UNIFORM64(pv[147]);
    const double P148 = PLOM_UNIFORM64(pv[148]);
    const double P149 = PLOM_UNIFORM64(pv[149]);
    const double P150 = PLOM_UNIFORM64(pv[150]);
    const double P151 = PLOM_UNIFORM64(pv[151]);
    const double P152 = PLOM_UNIFORM64(pv[152]);
    const double P153 = PLOM_UNIFORM64(pv[153]);
    const double P154 = PLOM_UNIFORM64(pv[154]);
    const double P155 = PLOM_UNIFORM64(pv[155]);
    const double P156 = PLOM_UNIFORM64(pv[156]);
    const double P157 = PLOM_UNIFORM64(pv[157]);
    const double P158 = PLOM_UNIFORM64(pv[158]);
    const double P159 = PLOM_UNIFORM64(pv[159]);
    const double P160 = PLOM_UNIFORM64(pv[160]);
    const double P161 = PLOM_UNIFORM64(pv[161]);
    const double P162 = PLOM_UNIFORM64(pv[162]);
    const double P163 = PLOM_UNIFORM64(pv[163]);
    const double P164 = PLOM_UNIFORM64(pv[164]);
    const double P165 = PLOM_UNIFORM64(pv[165]);
    const double P166 = PLOM_UNIFORM64(pv[166]);
    const double P167 = PLOM_UNIFORM64(pv[167]);
    const double P168 = PLOM_UNIFORM64(pv[168]);
    const double P169 = PLOM_UNIFORM64(pv[169]);
    const double P170 = PLOM_UNIFORM64(pv[170]);
    const double P171 = PLOM_UNIFORM64(pv[171]);
    const double P172 = PLOM_UNIFORM64(pv[172]);
    const double P173 = PLOM_UNIFORM64(pv[173]);
    const double P174 = PLOM_UNIFORM64(pv[174]);
    const double P175 = PLOM_UNIFORM64(pv[175]);
    const double P176 = PLOM_UNIFORM64(pv[176]);
    const double P177 = PLOM_UNIFORM64(pv[177]);
    const double P178 = PLOM_UNIFORM64(pv[178]);
    const double P179 = PLOM_UNIFORM64(pv[179]);
    X[30] = 0.0; /* reset_inc */
    X[31] = 0.0; /* reset_inc */
    X[32] = 0.0; /* reset_inc */
    X[33] = 0.0; /* reset_inc */
    X[34] = 0.0; /* reset_inc */
    X[35] = 0.0; /* reset_inc */
    X[36] = 0.0; /* reset_inc */
    X[37] = 0.0; /* reset_inc */
    X[38] = 0.0; /* reset_inc */
    X[39] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double C4 = plom_cov_ld(cvk + 4);
        const double C5 = plom_cov_ld(cvk + 5);
        const double C6 = plom_cov_ld(cvk + 6);
        const double C7 = plom_cov_ld(cvk + 7);
        const double C8 = plom_cov_ld(cvk + 8);
        const double C9 = plom_cov_ld(cvk + 9);
        double n0_0 = 0.0;
        double n1_0 = 0.0;
        double n2_0 = 0.0;
        double n3_0 = 0.0;
        double n0_1 = 0.0;
        double n1_1 = 0.0;
        double n2_1 = 0.0;
        double n3_1 = 0.0;
        double n0_2 = 0.0;
        double n1_2 = 0.0;
        double n2_2 = 0.0;
        double n3_2 = 0.0;
        double n0_3 = 0.0;
        double n1_3 = 0.0;
        double n2_3 = 0.0;
        double n3_3 = 0.0;
        double n0_4 = 0.0;
        double n1_4 = 0.0;
        double n2_4 = 0.0;
        double n3_4 = 0.0;
        double n0_5 = 0.0;
        double n1_5 = 0.0;
        double n2_5 = 0.0;
        double n3_5 = 0.0;
        double n0_6 = 0.0;
        double n1_6 = 0.0;
        double n2_6 = 0.0;
        double n3_6 = 0.0;
        double n0_7 = 0.0;
        double n1_7 = 0.0;
        double n2_7 = 0.0;
        double n3_7 = 0.0;
        double n0_8 = 0.0;
        double n1_8 = 0.0;
        double n2_8 = 0.0;
        double n3_8 = 0.0;
        double n0_9 = 0.0;
        double n1_9 = 0.0;
        double n2_9 = 0.0;
        double n3_9 = 0.0;
        double stay0 = 0.0;
        double stay1 = 0.0;
        double stay2 = 0.0;
        double stay3 = 0.0;
        double stay4 = 0.0;
        double stay5 = 0.0;
        double stay6 = 0.0;
        double stay7 = 0.0;
        double stay8 = 0.0;
        double stay9 = 0.0;
        double stay10 = 0.0;
        double stay11 = 0.0;
        double stay12 = 0.0;
        double stay13 = 0.0;
        double stay14 = 0.0;
        double stay15 = 0.0;
        double stay16 = 0.0;
        double stay17 = 0.0;
        double stay18 = 0.0;
        double stay19 = 0.0;
        double stay20 = 0.0;
        double stay21 = 0.0;
        double stay22 = 0.0;
        double stay23 = 0.0;
        double stay24 = 0.0;
        double stay25 = 0.0;
        double stay26 = 0.0;
        double stay27 = 0.0;
        double stay28 = 0.0;
        double stay29 = 0.0;
        (void)n0_0; (void)n1_0; (void)n2_0; (void)n3_0; (void)n0_1; (void)n1_1; (void)n2_1; (void)n3_1; (void)n0_2; (void)n1_2; (void)n2_2; (void)n3_2; (void)n0_3; (void)n1_3; (void)n2_3; (void)n3_3; (void)n0_4; (void)n1_4; (void)n2_4; (void)n3_4; (void)n0_5; (void)n1_5; (void)n2_5; (void)n3_5; (void)n0_6; (void)n1_6; (void)n2_6; (void)n3_6; (void)n0_7; (void)n1_7; (void)n2_7; (void)n3_7; (void)n0_8; (void)n1_8; (void)n2_8; (void)n3_8; (void)n0_9; (void)n1_9; (void)n2_9; (void)n3_9;
        const double r0_0 = ((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])) / C0);
        const double r1_0 = P10;
        const double r2_0 = P20;
        const double r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        const double r0_1 = ((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])) / C1);
        const double r1_1 = P11;
        const double r2_1 = P21;
        const double r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        const double r0_2 = ((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])) / C2);
        const double r1_2 = P12;
        const double r2_2 = P22;
        const double r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        const double r0_3 = ((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])) / C3);
        const double r1_3 = P13;
        const double r2_3 = P23;
        const double r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        const double r0_4 = ((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])) / C4);
        const double r1_4 = P14;
        const double r2_4 = P24;
        const double r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        const double r0_5 = ((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])) / C5);
        const double r1_5 = P15;
        const double r2_5 = P25;
        const double r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        const double r0_6 = ((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])) / C6);
        const double r1_6 = P16;
        const double r2_6 = P26;
        const double r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        const double r0_7 = ((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])) / C7);
        const double r1_7 = P17;
        const double r2_7 = P27;
        const double r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        const double r0_8 = ((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])) / C8);
        const double r1_8 = P18;
        const double r2_8 = P28;
        const double r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        const double r0_9 = ((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])) / C9);
        const double r1_9 = P19;
        const double r2_9 = P29;
        const double r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        const double xi0_0 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_0 * xi0_0) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_0 * xi0_0) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[0]), p, cnt);
            n0_0 = cnt[0];
            stay0 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_0 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_0 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[10]), p, cnt);
            n1_0 = cnt[0];
            stay10 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_0 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_0 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[20]), p, cnt);
            n2_0 = cnt[0];
            stay20 = cnt[1];
        }
        n3_0 = plom_poisson_draw(ns, r3_0 * dt);
        const double xi0_1 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_1 * xi0_1) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_1 * xi0_1) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[1]), p, cnt);
            n0_1 = cnt[0];
            stay1 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_1 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_1 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[11]), p, cnt);
            n1_1 = cnt[0];
            stay11 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_1 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_1 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[21]), p, cnt);
            n2_1 = cnt[0];
            stay21 = cnt[1];
        }
        n3_1 = plom_poisson_draw(ns, r3_1 * dt);
        const double xi0_2 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_2 * xi0_2) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_2 * xi0_2) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[2]), p, cnt);
            n0_2 = cnt[0];
            stay2 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_2 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_2 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[12]), p, cnt);
            n1_2 = cnt[0];
            stay12 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_2 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_2 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[22]), p, cnt);
            n2_2 = cnt[0];
            stay22 = cnt[1];
        }
        n3_2 = plom_poisson_draw(ns, r3_2 * dt);
        const double xi0_3 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_3 * xi0_3) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_3 * xi0_3) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[3]), p, cnt);
            n0_3 = cnt[0];
            stay3 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_3 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_3 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[13]), p, cnt);
            n1_3 = cnt[0];
            stay13 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_3 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_3 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[23]), p, cnt);
            n2_3 = cnt[0];
            stay23 = cnt[1];
        }
        n3_3 = plom_poisson_draw(ns, r3_3 * dt);
        const double xi0_4 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_4 * xi0_4) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_4 * xi0_4) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[4]), p, cnt);
            n0_4 = cnt[0];
            stay4 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_4 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_4 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[14]), p, cnt);
            n1_4 = cnt[0];
            stay14 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_4 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_4 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[24]), p, cnt);
            n2_4 = cnt[0];
            stay24 = cnt[1];
        }
        n3_4 = plom_poisson_draw(ns, r3_4 * dt);
        const double xi0_5 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_5 * xi0_5) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_5 * xi0_5) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[5]), p, cnt);
            n0_5 = cnt[0];
            stay5 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_5 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_5 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[15]), p, cnt);
            n1_5 = cnt[0];
            stay15 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_5 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_5 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[25]), p, cnt);
            n2_5 = cnt[0];
            stay25 = cnt[1];
        }
        n3_5 = plom_poisson_draw(ns, r3_5 * dt);
        const double xi0_6 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_6 * xi0_6) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_6 * xi0_6) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[6]), p, cnt);
            n0_6 = cnt[0];
            stay6 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_6 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_6 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[16]), p, cnt);
            n1_6 = cnt[0];
            stay16 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_6 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_6 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[26]), p, cnt);
            n2_6 = cnt[0];
            stay26 = cnt[1];
        }
        n3_6 = plom_poisson_draw(ns, r3_6 * dt);
        const double xi0_7 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_7 * xi0_7) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_7 * xi0_7) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[7]), p, cnt);
            n0_7 = cnt[0];
            stay7 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_7 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_7 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[17]), p, cnt);
            n1_7 = cnt[0];
            stay17 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_7 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_7 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[27]), p, cnt);
            n2_7 = cnt[0];
            stay27 = cnt[1];
        }
        n3_7 = plom_poisson_draw(ns, r3_7 * dt);
        const double xi0_8 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_8 * xi0_8) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_8 * xi0_8) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[8]), p, cnt);
            n0_8 = cnt[0];
            stay8 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_8 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_8 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[18]), p, cnt);
            n1_8 = cnt[0];
            stay18 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_8 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_8 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[28]), p, cnt);
            n2_8 = cnt[0];
            stay28 = cnt[1];
        }
        n3_8 = plom_poisson_draw(ns, r3_8 * dt);
        const double xi0_9 = 1.0;
        {
            double p[2], cnt[2];
            const double sum = (r0_9 * xi0_9) * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * (((r0_9 * xi0_9) * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[9]), p, cnt);
            n0_9 = cnt[0];
            stay9 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r1_9 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r1_9 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[19]), p, cnt);
            n1_9 = cnt[0];
            stay19 = cnt[1];
        }
        {
            double p[2], cnt[2];
            const double sum = r2_9 * dt;
            if (sum > 0.0) {
                const double one_minus_exp_sum = 1.0 - plom_exp(-sum);
                p[0] = one_minus_exp_sum * ((r2_9 * dt) / sum);
                p[1] = 1.0 - p[0];
            } else {
                p[0] = 0.0;
                p[1] = 1.0;
            }
            plom_multinomial(ns, 2, plom_floor(X[29]), p, cnt);
            n2_9 = cnt[0];
            stay29 = cnt[1];
        }
        n3_9 = plom_poisson_draw(ns, r3_9 * dt);
        X[0] = stay0 + n3_0;
        X[10] = stay10 + n0_0;
        X[20] = stay20 + n1_0;
        X[1] = stay1 + n3_1;
        X[11] = stay11 + n0_1;
        X[21] = stay21 + n1_1;
        X[2] = stay2 + n3_2;
        X[12] = stay12 + n0_2;
        X[22] = stay22 + n1_2;
        X[3] = stay3 + n3_3;
        X[13] = stay13 + n0_3;
        X[23] = stay23 + n1_3;
        X[4] = stay4 + n3_4;
        X[14] = stay14 + n0_4;
        X[24] = stay24 + n1_4;
        X[5] = stay5 + n3_5;
        X[15] = stay15 + n0_5;
        X[25] = stay25 + n1_5;
        X[6] = stay6 + n3_6;
        X[16] = stay16 + n0_6;
        X[26] = stay26 + n1_6;
        X[7] = stay7 + n3_7;
        X[17] = stay17 + n0_7;
        X[27] = stay27 + n1_7;
        X[8] = stay8 + n3_8;
        X[18] = stay18 + n0_8;
        X[28] = stay28 + n1_8;
        X[9] = stay9 + n3_9;
        X[19] = stay19 + n0_9;
        X[29] = stay29 + n1_9;
        X[30] = X[30] + (n1_0);
        X[31] = X[31] + (n1_1);
        X[32] = X[32] + (n1_2);
        X[33] = X[33] + (n1_3);
        X[34] = X[34] + (n1_4);
        X[35] = X[35] + (n1_5);
        X[36] = X[36] + (n1_6);
        X[37] = X[37] + (n1_7);
        X[38] = X[38] + (n1_8);
        X[39] = X[39] + (n1_9);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_ode_rhs_pp(double t, const double *X, const double *pv, const double *lp, const double *cv, const plom_spline_set *sp, double *f)
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
    const double P16 = lp[16];
    const double P17 = lp[17];
    const double P18 = lp[18];
    const double P19 = lp[19];
    const double P20 = lp[20];
    const double P21 = lp[21];
    const double P22 = lp[22];
    const double P23 = lp[23];
    const double P24 = lp[24];
    const double P25 = lp[25];
    const double P26 = lp[26];
    const double P27 = lp[27];
    const double P28 = lp[28];
    const double P29 = lp[29];
    const double P30 = lp[30];
    const double P31 = lp[31];
    const double P32 = lp[32];
    const double P33 = lp[33];
    const double P34 = lp[34];
    const double P35 = lp[35];
    const double P36 = lp[36];
    const double P37 = lp[37];
    const double P38 = lp[38];
    const double P39 = lp[39];
    const double P80 = PLOM_UNIFORM64(pv[80]);
    const double P81 = PLOM_UNIFORM64(pv[81]);
    const double P82 = PLOM_UNIFORM64(pv[82]);
    const double P83 = PLOM_UNIFORM64(pv[83]);
    const double P84 = PLOM_UNIFORM64(pv[84]);
    const double P85 = PLOM_UNIFORM64(pv[85]);
    const double P86 = PLOM_UNIFORM64(pv[86]);
    const double P87 = PLOM_UNIFORM64(pv[87]);
    const double P88 = PLOM_UNIFORM64(pv[88]);
    const double P89 = PLOM_UNIFORM64(pv[89]);
    const double P90 = PLOM_UNIFORM64(pv[90]);
    const double P91 = PLOM_UNIFORM64(pv[91]);
    const double P92 = PLOM_UNIFORM64(pv[92]);
    const double P93 = PLOM_UNIFORM64(pv[93]);
    const double P94 = PLOM_UNIFORM64(pv[94]);
    const double P95 = PLOM_UNIFORM64(pv[95]);
    const double P96 = PLOM_UNIFORM64(pv[96]);
    const double P97 = PLOM_UNIFORM64(pv[97]);
    const double P98 = PLOM_UNIFORM64(pv[98]);
    const double P99 = PLOM_UNIFORM64(pv[99]);
    const double P100 = PLOM_UNIFORM64(pv[100]);
    const double P101 = PLOM_UNIFORM64(pv[101]);
    const double P102 = PLOM_UNIFORM64(pv[102]);
    const double P103 = PLOM_UNIFORM64(pv[103]);
    const double P104 = PLOM_UNIFORM64(pv[104]);
    const double P105 = PLOM_UNIFORM64(pv[105]);
    const double P106 = PLOM_UNIFORM64(pv[106]);
    const double P107 = PLOM_UNIFORM64(pv[107]);
    const double P108 = PLOM_UNIFORM64(pv[108]);
    const double P109 = PLOM_UNIFORM64(pv[109]);
    const double P110 = PLOM_UNIFORM64(pv[110]);
    const double P111 = PLOM_UNIFORM64(pv[111]);
    const double P112 = PLOM_UNIFORM64(pv[112]);
    const double P113 = PLOM_UNIFORM64(pv[113]);
    const double P114 = PLOM_UNIFORM64(pv[114]);
    const double P115 = PLOM_UNIFORM64(pv[115]);
    const double P116 = PLOM_UNIFORM64(pv[116]);
    const double P117 = PLOM_UNIFORM64(pv[117]);
    const double P118 = PLOM_UNIFORM64(pv[118]);
    const double P119 = PLOM_UNIFORM64(pv[119]);
    const double P120 = PLOM_UNIFORM64(pv[120]);
    const double P121 = PLOM_UNIFORM64(pv[121]);
    const double P122 = PLOM_UNIFORM64(pv[122]);
    const double P123 = PLOM_UNIFORM64(pv[123]);
    const double P124 = PLOM_UNIFORM64(pv[124]);
    const double P125 = PLOM_UNIFORM64(pv[125]);
    const double P126 = PLOM_UNIFORM64(pv[126]);
    const double P127 = PLOM_UNIFORM64(pv[127]);
    const double P128 = PLOM_UNIFORM64(pv[128]);
    const double P129 = PLOM_UNIFORM64(pv[129]);
    const double P130 = PLOM_UNIFORM64(pv[130]);
    const double P131 = PLOM_UNIFORM64(pv[131]);
    const double P132 = PLOM_UNIFORM64(pv[132]);
    const double P133 = PLOM_UNIFORM64(pv[133]);
    const double P134 = PLOM_UNIFORM64(pv[134]);
    const double P135 = PLOM_UNIFORM64(pv[135]);
    const double P136 = PLOM_UNIFORM64(pv[136]);
    const double P137 = PLOM_UNIFORM64(pv[137]);
    const double P138 = PLOM_UNIFORM64(pv[138]);
    const double P139 = PLOM_UNIFORM64(pv[139]);
    const double P140 = PLOM_UNIFORM64(pv[140]);
    const double P141 = PLOM_UNIFORM64(pv[141]);
    const double P142 = PLOM_UNIFORM64(pv[142]);
    const double P143 = PLOM_UNIFORM64(pv[143]);
    const double P144 = PLOM_UNIFORM64(pv[144]);
    const double P145 = PLOM_UNIFORM64(pv[145]);
    const double P146 = PLOM_UNIFORM64(pv[146]);
    const double P147 = PLOM_UNIFORM64(pv[147]);
    const double P148 = PLOM_UNIFORM64(pv[148]);
    const double P149 = PLOM_UNIFORM64(pv[149]);
    const double P150 = PLOM_UNIFORM64(pv[150]);
    const double P151 = PLOM_UNIFORM64(pv[151]);
    const double P152 = PLOM_UNIFORM64(pv[152]);
    const double P153 = PLOM_UNIFORM64(pv[153]);
    const double P154 = PLOM_UNIFORM64(pv[154]);
    const double P155 = PLOM_UNIFORM64(pv[155]);
    const double P156 = PLOM_UNIFORM64(pv[156]);
    const double P157 = PLOM_UNIFORM64(pv[157]);
    const double P158 = PLOM_UNIFORM64(pv[158]);
    const double P159 = PLOM_UNIFORM64(pv[159]);
    const double P160 = PLOM_UNIFORM64(pv[160]);
    const double P161 = PLOM_UNIFORM64(pv[161]);
    const double P162 = PLOM_UNIFORM64(pv[162]);
    const double P163 = PLOM_UNIFORM64(pv[163]);
    const double P164 = PLOM_UNIFORM64(pv[164]);
    const double P165 = PLOM_UNIFORM64(pv[165]);
    const double P166 = PLOM_UNIFORM64(pv[166]);
    const double P167 = PLOM_UNIFORM64(pv[167]);
    const double P168 = PLOM_UNIFORM64(pv[168]);
    const double P169 = PLOM_UNIFORM64(pv[169]);
    const double P170 = PLOM_UNIFORM64(pv[170]);
    const double P171 = PLOM_UNIFORM64(pv[171]);
    const double P172 = PLOM_UNIFORM64(pv[172]);
    const double P173 = PLOM_UNIFORM64(pv[173]);
    const double P174 = PLOM_UNIFORM64(pv[174]);
    const double P175 = PLOM_UNIFORM64(pv[175]);
    const double P176 = PLOM_UNIFORM64(pv[176]);
    const double P177 = PLOM_UNIFORM64(pv[177]);
    const double P178 = PLOM_UNIFORM64(pv[178]);
    const double P179 = PLOM_UNIFORM64(pv[179]);
    const double C0 = sp ? plom_spline_at(sp, 0, t) : plom_cov_ld(cv + 0);
    const double C1 = sp ? plom_spline_at(sp, 1, t) : plom_cov_ld(cv + 1);
    const double C2 = sp ? plom_spline_at(sp, 2, t) : plom_cov_ld(cv + 2);
    const double C3 = sp ? plom_spline_at(sp, 3, t) : plom_cov_ld(cv + 3);
    const double C4 = sp ? plom_spline_at(sp, 4, t) : plom_cov_ld(cv + 4);
    const double C5 = sp ? plom_spline_at(sp, 5, t) : plom_cov_ld(cv + 5);
    const double C6 = sp ? plom_spline_at(sp, 6, t) : plom_cov_ld(cv + 6);
    const double C7 = sp ? plom_spline_at(sp, 7, t) : plom_cov_ld(cv + 7);
    const double C8 = sp ? plom_spline_at(sp, 8, t) : plom_cov_ld(cv + 8);
    const double C9 = sp ? plom_spline_at(sp, 9, t) : plom_cov_ld(cv + 9);
    const double r0_0 = ((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])) / C0);
    const double r1_0 = P10;
    const double r2_0 = P20;
    const double r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
    const double r0_1 = ((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])) / C1);
    const double r1_1 = P11;
    const double r2_1 = P21;
    const double r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
    const double r0_2 = ((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])) / C2);
    const double r1_2 = P12;
    const double r2_2 = P22;
    const double r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
    const double r0_3 = ((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])) / C3);
    const double r1_3 = P13;
    const double r2_3 = P23;
    const double r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
    const double r0_4 = ((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])) / C4);
    const double r1_4 = P14;
    const double r2_4 = P24;
    const double r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
    const double r0_5 = ((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])) / C5);
    const double r1_5 = P15;
    const double r2_5 = P25;
    const double r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
    const double r0_6 = ((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])) / C6);
    const double r1_6 = P16;
    const double r2_6 = P26;
    const double r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
    const double r0_7 = ((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])) / C7);
    const double r1_7 = P17;
    const double r2_7 = P27;
    const double r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
    const double r0_8 = ((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])) / C8);
    const double r1_8 = P18;
    const double r2_8 = P28;
    const double r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
    const double r0_9 = ((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])) / C9);
    const double r1_9 = P19;
    const double r2_9 = P29;
    const double r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
    f[0] =  - (r0_0*X[0]) + (r3_0);
    f[10] =  - (r1_0*X[10]) + (r0_0*X[0]);
    f[20] =  - (r2_0*X[20]) + (r1_0*X[10]);
    f[1] =  - (r0_1*X[1]) + (r3_1);
    f[11] =  - (r1_1*X[11]) + (r0_1*X[1]);
    f[21] =  - (r2_1*X[21]) + (r1_1*X[11]);
    f[2] =  - (r0_2*X[2]) + (r3_2);
    f[12] =  - (r1_2*X[12]) + (r0_2*X[2]);
    f[22] =  - (r2_2*X[22]) + (r1_2*X[12]);
    f[3] =  - (r0_3*X[3]) + (r3_3);
    f[13] =  - (r1_3*X[13]) + (r0_3*X[3]);
    f[23] =  - (r2_3*X[23]) + (r1_3*X[13]);
    f[4] =  - (r0_4*X[4]) + (r3_4);
    f[14] =  - (r1_4*X[14]) + (r0_4*X[4]);
    f[24] =  - (r2_4*X[24]) + (r1_4*X[14]);
    f[5] =  - (r0_5*X[5]) + (r3_5);
    f[15] =  - (r1_5*X[15]) + (r0_5*X[5]);
    f[25] =  - (r2_5*X[25]) + (r1_5*X[15]);
    f[6] =  - (r0_6*X[6]) + (r3_6);
    f[16] =  - (r1_6*X[16]) + (r0_6*X[6]);
    f[26] =  - (r2_6*X[26]) + (r1_6*X[16]);
    f[7] =  - (r0_7*X[7]) + (r3_7);
    f[17] =  - (r1_7*X[17]) + (r0_7*X[7]);
    f[27] =  - (r2_7*X[27]) + (r1_7*X[17]);
    f[8] =  - (r0_8*X[8]) + (r3_8);
    f[18] =  - (r1_8*X[18]) + (r0_8*X[8]);
    f[28] =  - (r2_8*X[28]) + (r1_8*X[18]);
    f[9] =  - (r0_9*X[9]) + (r3_9);
    f[19] =  - (r1_9*X[19]) + (r0_9*X[9]);
    f[29] =  - (r2_9*X[29]) + (r1_9*X[19]);
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_0*X[10]);
        f[30] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_1*X[11]);
        f[31] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_2*X[12]);
        f[32] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_3*X[13]);
        f[33] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_4*X[14]);
        f[34] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_5*X[15]);
        f[35] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_6*X[16]);
        f[36] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_7*X[17]);
        f[37] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_8*X[18]);
        f[38] = sum_inc;
    }
    {
        double sum_inc = 0.0;
        sum_inc +=  + (r1_9*X[19]);
        f[39] = sum_inc;
    }
}

#define PLOM_ODE_FN plom_seir10_ode_integrate_pp
#define PLOM_ODE_RHS plom_seir10_ode_rhs_pp
#define PLOM_ODE_N 40
#define PLOM_ODE_QUAL static __device__ inline __attribute__((always_inline))
#include "plom_ode_impl.h"

static __device__ inline __attribute__((always_inline)) int plom_seir10_interval_ode_pp_f64(double *X, const double *pv, const double *lp, const double *cv, const plom_spline_set *sp, double t0, double t1, double dt, double eps_abs, double eps_rel)
{
    X[30] = 0.0; /* reset_inc */
    X[31] = 0.0; /* reset_inc */
    X[32] = 0.0; /* reset_inc */
    X[33] = 0.0; /* reset_inc */
    X[34] = 0.0; /* reset_inc */
    X[35] = 0.0; /* reset_inc */
    X[36] = 0.0; /* reset_inc */
    X[37] = 0.0; /* reset_inc */
    X[38] = 0.0; /* reset_inc */
    X[39] = 0.0; /* reset_inc */
    return plom_seir10_ode_integrate_pp(X, pv, lp, cv, sp, t0, t1, dt, eps_abs, eps_rel);
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_interval_full_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
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
    const double P20 = lp[20];
    const double P21 = lp[21];
    const double P22 = lp[22];
    const double P23 = lp[23];
    const double P24 = lp[24];
    const double P25 = lp[25];
    const double P26 = lp[26];
    const double P27 = lp[27];
    const double P28 = lp[28];
    const double P29 = lp[29];
    const double P30 = lp[30];
    const double P31 = lp[31];
    const double P32 = lp[32];
    const double P33 = lp[33];
    const double P34 = lp[34];
    const double P35 = lp[35];
    const double P36 = lp[36];
    const double P37 = lp[37];
    const double P38 = lp[38];
    const double P39 = lp[39];
    const double P40 = lp[40];
    const double P41 = lp[41];
    const double P42 = lp[42];
    const double P43 = lp[43];
    const double P44 = lp[44];
    const double P45 = lp[45];
    const double P46 = lp[46];
    const double P47 = lp[47];
    const double P48 = lp[48];
    const double P49 = lp[49];
    const double P80 = PLOM_UNIFORM64(pv[80]);
    const double P81 = PLOM_UNIFORM64(pv[81]);
    const double P82 = PLOM_UNIFORM64(pv[82]);
    const double P83 = PLOM_UNIFORM64(pv[83]);
    const double P84 = PLOM_UNIFORM64(pv[84]);
    const double P85 = PLOM_UNIFORM64(pv[85]);
    const double P86 = PLOM_UNIFORM64(pv[86]);
    const double P87 = PLOM_UNIFORM64(pv[87]);
    const double P88 = PLOM_UNIFORM64(pv[88]);
    const double P89 = PLOM_UNIFORM64(pv[89]);
    const double P90 = PLOM_UNIFORM64(pv[90]);
    const double P91 = PLOM_UNIFORM64(pv[91]);
    const double P92 = PLOM_UNIFORM64(pv[92]);
    const double P93 = PLOM_UNIFORM64(pv[93]);
    const double P94 = PLOM_UNIFORM64(pv[94]);
    const double P95 = PLOM_UNIFORM64(pv[95]);
    const double P96 = PLOM_UNIFORM64(pv[96]);
    const double P97 = PLOM_UNIFORM64(pv[97]);
    const double P98 = PLOM_UNIFORM64(pv[98]);
    const double P99 = PLOM_UNIFORM64(pv[99]);
    const double P100 = PLOM_UNIFORM64(pv[100]);
    const double P101 = PLOM_UNIFORM64(pv[101]);
    const double P102 = PLOM_UNIFORM64(pv[102]);
    const double P103 = PLOM_UNIFORM64(pv[103]);
    const double P104 = PLOM_UNIFORM64(pv[104]);
    const double P105 = PLOM_UNIFORM64(pv[105]);
    const double P106 = PLOM_UNIFORM64(pv[106]);
    const double P107 = PLOM_UNIFORM64(pv[107]);
    const double P108 = PLOM_UNIFORM64(pv[108]);
    const double P109 = PLOM_UNIFORM64(pv[109]);
    const double P110 = PLOM_UNIFORM64(pv[110]);
    const double P111 = PLOM_UNIFORM64(pv[111]);
    const double P112 = PLOM_UNIFORM64(pv[112]);
    const double P113 = PLOM_UNIFORM64(pv[113]);
    const double P114 = PLOM_UNIFORM64(pv[114]);
    const double P115 = PLOM_UNIFORM64(pv[115]);
    const double P116 = PLOM_UNIFORM64(pv[116]);
    const double P117 = PLOM_UNIFORM64(pv[117]);
    const double P118 = PLOM_UNIFORM64(pv[118]);
    const double P119 = PLOM_UNIFORM64(pv[119]);
    const double P120 = PLOM_UNIFORM64(pv[120]);
    const double P121 = PLOM_UNIFORM64(pv[121]);
    const double P122 = PLOM_UNIFORM64(pv[122]);
    const double P123 = PLOM_UNIFORM64(pv[123]);
    const double P124 = PLOM_UNIFORM64(pv[124]);
    const double P125 = PLOM_UNIFORM64(pv[125]);
    const double P126 = PLOM_UNIFORM64(pv[126]);
    const double P127 = PLOM_UNIFORM64(pv[127]);
    const double P128 = PLOM_UNIFORM64(pv[128]);
    const double P129 = PLOM_UNIFORM64(pv[129]);
    const double P130 = PLOM_UNIFORM64(pv[130]);
    const double P131 = PLOM_UNIFORM64(pv[131]);
    const double P132 = PLOM_UNIFORM64(pv[132]);
    const double P133 = PLOM_UNIFORM64(pv[133]);
    const double P134 = PLOM_UNIFORM64(pv[134]);
    const double P135 = PLOM_UNIFORM64(pv[135]);
    const double P136 = PLOM_UNIFORM64(pv[136]);
    const double P137 = PLOM_UNIFORM64(pv[137]);
    const double P138 = PLOM_UNIFORM64(pv[138]);
    const double P139 = PLOM_UNIFORM64(pv[139]);
    const double P140 = PLOM_UNIFORM64(pv[140]);
    const double P141 = PLOM_UNIFORM64(pv[141]);
    const double P142 = PLOM_UNIFORM64(pv[142]);
    const double P143 = PLOM_UNIFORM64(pv[143]);
    const double P144 = PLOM_UNIFORM64(pv[144]);
    const double P145 = PLOM_UNIFORM64(pv[145]);
    const double P146 = PLOM_UNIFORM64(pv[146]);
    const double P147 = PLOM_UNIFORM64(pv[147]);
    const double P148 = PLOM_UNIFORM64(pv[148]);
    const double P149 = PLOM_UNIFORM64(pv[149]);
    const double P150 = PLOM_UNIFORM64(pv[150]);
    const double P151 = PLOM_UNIFORM64(pv[151]);
    const double P152 = PLOM_UNIFORM64(pv[152]);
    const double P153 = PLOM_UNIFORM64(pv[153]);
    const double P154 = PLOM_UNIFORM64(pv[154]);
    const double P155 = PLOM_UNIFORM64(pv[155]);
    const double P156 = PLOM_UNIFORM64(pv[156]);
    const double P157 = PLOM_UNIFORM64(pv[157]);
    const double P158 = PLOM_UNIFORM64(pv[158]);
    const double P159 = PLOM_UNIFORM64(pv[159]);
    const double P160 = PLOM_UNIFORM64(pv[160]);
    const double P161 = PLOM_UNIFORM64(pv[161]);
    const double P162 = PLOM_UNIFORM64(pv[162]);
    const double P163 = PLOM_UNIFORM64(pv[163]);
    const double P164 = PLOM_UNIFORM64(pv[164]);
    const double P165 = PLOM_UNIFORM64(pv[165]);
    const double P166 = PLOM_UNIFORM64(pv[166]);
    const double P167 = PLOM_UNIFORM64(pv[167]);
    const double P168 = PLOM_UNIFORM64(pv[168]);
    const double P169 = PLOM_UNIFORM64(pv[169]);
    const double P170 = PLOM_UNIFORM64(pv[170]);
    const double P171 = PLOM_UNIFORM64(pv[171]);
    const double P172 = PLOM_UNIFORM64(pv[172]);
    const double P173 = PLOM_UNIFORM64(pv[173]);
    const double P174 = PLOM_UNIFORM64(pv[174]);
    const double P175 = PLOM_UNIFORM64(pv[175]);
    const double P176 = PLOM_UNIFORM64(pv[176]);
    const double P177 = PLOM_UNIFORM64(pv[177]);
    const double P178 = PLOM_UNIFORM64(pv[178]);
    const double P179 = PLOM_UNIFORM64(pv[179]);
    X[30] = 0.0; /* reset_inc */
    X[31] = 0.0; /* reset_inc */
    X[32] = 0.0; /* reset_inc */
    X[33] = 0.0; /* reset_inc */
    X[34] = 0.0; /* reset_inc */
    X[35] = 0.0; /* reset_inc */
    X[36] = 0.0; /* reset_inc */
    X[37] = 0.0; /* reset_inc */
    X[38] = 0.0; /* reset_inc */
    X[39] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double C4 = plom_cov_ld(cvk + 4);
        const double C5 = plom_cov_ld(cvk + 5);
        const double C6 = plom_cov_ld(cvk + 6);
        const double C7 = plom_cov_ld(cvk + 7);
        const double C8 = plom_cov_ld(cvk + 8);
        const double C9 = plom_cov_ld(cvk + 9);
        double s30 = 0.0;
        double s31 = 0.0;
        double s32 = 0.0;
        double s33 = 0.0;
        double s34 = 0.0;
        double s35 = 0.0;
        double s36 = 0.0;
        double s37 = 0.0;
        double s38 = 0.0;
        double s39 = 0.0;
        const double r0_0 = ((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])) / C0);
        const double r1_0 = P10;
        const double r2_0 = P20;
        const double r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        double za0, zb0;
        plom_normal_pair64(ns, &za0, &zb0);
        const double d0_0 = sqdt * za0;
        const double d1_0 = sqdt * zb0;
        double za1, zb1;
        plom_normal_pair64(ns, &za1, &zb1);
        const double d2_0 = sqdt * za1;
        const double d3_0 = sqdt * zb1;
        double za2, zb2;
        plom_normal_pair64(ns, &za2, &zb2);
        const double w0_0 = sqdt * za2;
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r3_0))*dtr + - plom_sqrt_rate((r0_0*X[0]))*d0_0+ plom_sqrt_rate((r3_0))*d3_0 + - (r0_0*X[0])*P40*w0_0;
        const double f10 = X[10] + ( - (r1_0*X[10]) + (r0_0*X[0]))*dtr + - plom_sqrt_rate((r1_0*X[10]))*d1_0+ plom_sqrt_rate((r0_0*X[0]))*d0_0 + + (r0_0*X[0])*P40*w0_0;
        const double f20 = X[20] + ( - (r2_0*X[20]) + (r1_0*X[10]))*dtr + - plom_sqrt_rate((r2_0*X[20]))*d2_0+ plom_sqrt_rate((r1_0*X[10]))*d1_0;
        s30 += ( + (r1_0*X[10]))*dtr + + plom_sqrt_rate((r1_0*X[10]))*d1_0;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        const double r0_1 = ((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])) / C1);
        const double r1_1 = P11;
        const double r2_1 = P21;
        const double r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        const double d0_1 = sqdt * zb2;
        double za3, zb3;
        plom_normal_pair64(ns, &za3, &zb3);
        const double d1_1 = sqdt * za3;
        const double d2_1 = sqdt * zb3;
        double za4, zb4;
        plom_normal_pair64(ns, &za4, &zb4);
        const double d3_1 = sqdt * za4;
        const double w0_1 = sqdt * zb4;
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r3_1))*dtr + - plom_sqrt_rate((r0_1*X[1]))*d0_1+ plom_sqrt_rate((r3_1))*d3_1 + - (r0_1*X[1])*P41*w0_1;
        const double f11 = X[11] + ( - (r1_1*X[11]) + (r0_1*X[1]))*dtr + - plom_sqrt_rate((r1_1*X[11]))*d1_1+ plom_sqrt_rate((r0_1*X[1]))*d0_1 + + (r0_1*X[1])*P41*w0_1;
        const double f21 = X[21] + ( - (r2_1*X[21]) + (r1_1*X[11]))*dtr + - plom_sqrt_rate((r2_1*X[21]))*d2_1+ plom_sqrt_rate((r1_1*X[11]))*d1_1;
        s31 += ( + (r1_1*X[11]))*dtr + + plom_sqrt_rate((r1_1*X[11]))*d1_1;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        const double r0_2 = ((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])) / C2);
        const double r1_2 = P12;
        const double r2_2 = P22;
        const double r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        double za5, zb5;
        plom_normal_pair64(ns, &za5, &zb5);
        const double d0_2 = sqdt * za5;
        const double d1_2 = sqdt * zb5;
        double za6, zb6;
        plom_normal_pair64(ns, &za6, &zb6);
        const double d2_2 = sqdt * za6;
        const double d3_2 = sqdt * zb6;
        double za7, zb7;
        plom_normal_pair64(ns, &za7, &zb7);
        const double w0_2 = sqdt * za7;
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r3_2))*dtr + - plom_sqrt_rate((r0_2*X[2]))*d0_2+ plom_sqrt_rate((r3_2))*d3_2 + - (r0_2*X[2])*P42*w0_2;
        const double f12 = X[12] + ( - (r1_2*X[12]) + (r0_2*X[2]))*dtr + - plom_sqrt_rate((r1_2*X[12]))*d1_2+ plom_sqrt_rate((r0_2*X[2]))*d0_2 + + (r0_2*X[2])*P42*w0_2;
        const double f22 = X[22] + ( - (r2_2*X[22]) + (r1_2*X[12]))*dtr + - plom_sqrt_rate((r2_2*X[22]))*d2_2+ plom_sqrt_rate((r1_2*X[12]))*d1_2;
        s32 += ( + (r1_2*X[12]))*dtr + + plom_sqrt_rate((r1_2*X[12]))*d1_2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        X[12] = (f12 < 0.0) ? 0.0 : f12;
        const double r0_3 = ((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])) / C3);
        const double r1_3 = P13;
        const double r2_3 = P23;
        const double r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        const double d0_3 = sqdt * zb7;
        double za8, zb8;
        plom_normal_pair64(ns, &za8, &zb8);
        const double d1_3 = sqdt * za8;
        const double d2_3 = sqdt * zb8;
        double za9, zb9;
        plom_normal_pair64(ns, &za9, &zb9);
        const double d3_3 = sqdt * za9;
        const double w0_3 = sqdt * zb9;
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r3_3))*dtr + - plom_sqrt_rate((r0_3*X[3]))*d0_3+ plom_sqrt_rate((r3_3))*d3_3 + - (r0_3*X[3])*P43*w0_3;
        const double f13 = X[13] + ( - (r1_3*X[13]) + (r0_3*X[3]))*dtr + - plom_sqrt_rate((r1_3*X[13]))*d1_3+ plom_sqrt_rate((r0_3*X[3]))*d0_3 + + (r0_3*X[3])*P43*w0_3;
        const double f23 = X[23] + ( - (r2_3*X[23]) + (r1_3*X[13]))*dtr + - plom_sqrt_rate((r2_3*X[23]))*d2_3+ plom_sqrt_rate((r1_3*X[13]))*d1_3;
        s33 += ( + (r1_3*X[13]))*dtr + + plom_sqrt_rate((r1_3*X[13]))*d1_3;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[13] = (f13 < 0.0) ? 0.0 : f13;
        const double r0_4 = ((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])) / C4);
        const double r1_4 = P14;
        const double r2_4 = P24;
        const double r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        double za10, zb10;
        plom_normal_pair64(ns, &za10, &zb10);
        const double d0_4 = sqdt * za10;
        const double d1_4 = sqdt * zb10;
        double za11, zb11;
        plom_normal_pair64(ns, &za11, &zb11);
        const double d2_4 = sqdt * za11;
        const double d3_4 = sqdt * zb11;
        double za12, zb12;
        plom_normal_pair64(ns, &za12, &zb12);
        const double w0_4 = sqdt * za12;
        const double f4 = X[4] + ( - (r0_4*X[4]) + (r3_4))*dtr + - plom_sqrt_rate((r0_4*X[4]))*d0_4+ plom_sqrt_rate((r3_4))*d3_4 + - (r0_4*X[4])*P44*w0_4;
        const double f14 = X[14] + ( - (r1_4*X[14]) + (r0_4*X[4]))*dtr + - plom_sqrt_rate((r1_4*X[14]))*d1_4+ plom_sqrt_rate((r0_4*X[4]))*d0_4 + + (r0_4*X[4])*P44*w0_4;
        const double f24 = X[24] + ( - (r2_4*X[24]) + (r1_4*X[14]))*dtr + - plom_sqrt_rate((r2_4*X[24]))*d2_4+ plom_sqrt_rate((r1_4*X[14]))*d1_4;
        s34 += ( + (r1_4*X[14]))*dtr + + plom_sqrt_rate((r1_4*X[14]))*d1_4;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[14] = (f14 < 0.0) ? 0.0 : f14;
        const double r0_5 = ((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])) / C5);
        const double r1_5 = P15;
        const double r2_5 = P25;
        const double r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        const double d0_5 = sqdt * zb12;
        double za13, zb13;
        plom_normal_pair64(ns, &za13, &zb13);
        const double d1_5 = sqdt * za13;
        const double d2_5 = sqdt * zb13;
        double za14, zb14;
        plom_normal_pair64(ns, &za14, &zb14);
        const double d3_5 = sqdt * za14;
        const double w0_5 = sqdt * zb14;
        const double f5 = X[5] + ( - (r0_5*X[5]) + (r3_5))*dtr + - plom_sqrt_rate((r0_5*X[5]))*d0_5+ plom_sqrt_rate((r3_5))*d3_5 + - (r0_5*X[5])*P45*w0_5;
        const double f15 = X[15] + ( - (r1_5*X[15]) + (r0_5*X[5]))*dtr + - plom_sqrt_rate((r1_5*X[15]))*d1_5+ plom_sqrt_rate((r0_5*X[5]))*d0_5 + + (r0_5*X[5])*P45*w0_5;
        const double f25 = X[25] + ( - (r2_5*X[25]) + (r1_5*X[15]))*dtr + - plom_sqrt_rate((r2_5*X[25]))*d2_5+ plom_sqrt_rate((r1_5*X[15]))*d1_5;
        s35 += ( + (r1_5*X[15]))*dtr + + plom_sqrt_rate((r1_5*X[15]))*d1_5;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[15] = (f15 < 0.0) ? 0.0 : f15;
        const double r0_6 = ((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])) / C6);
        const double r1_6 = P16;
        const double r2_6 = P26;
        const double r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        double za15, zb15;
        plom_normal_pair64(ns, &za15, &zb15);
        const double d0_6 = sqdt * za15;
        const double d1_6 = sqdt * zb15;
        double za16, zb16;
        plom_normal_pair64(ns, &za16, &zb16);
        const double d2_6 = sqdt * za16;
        const double d3_6 = sqdt * zb16;
        double za17, zb17;
        plom_normal_pair64(ns, &za17, &zb17);
        const double w0_6 = sqdt * za17;
        const double f6 = X[6] + ( - (r0_6*X[6]) + (r3_6))*dtr + - plom_sqrt_rate((r0_6*X[6]))*d0_6+ plom_sqrt_rate((r3_6))*d3_6 + - (r0_6*X[6])*P46*w0_6;
        const double f16 = X[16] + ( - (r1_6*X[16]) + (r0_6*X[6]))*dtr + - plom_sqrt_rate((r1_6*X[16]))*d1_6+ plom_sqrt_rate((r0_6*X[6]))*d0_6 + + (r0_6*X[6])*P46*w0_6;
        const double f26 = X[26] + ( - (r2_6*X[26]) + (r1_6*X[16]))*dtr + - plom_sqrt_rate((r2_6*X[26]))*d2_6+ plom_sqrt_rate((r1_6*X[16]))*d1_6;
        s36 += ( + (r1_6*X[16]))*dtr + + plom_sqrt_rate((r1_6*X[16]))*d1_6;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[16] = (f16 < 0.0) ? 0.0 : f16;
        const double r0_7 = ((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])) / C7);
        const double r1_7 = P17;
        const double r2_7 = P27;
        const double r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        const double d0_7 = sqdt * zb17;
        double za18, zb18;
        plom_normal_pair64(ns, &za18, &zb18);
        const double d1_7 = sqdt * za18;
        const double d2_7 = sqdt * zb18;
        double za19, zb19;
        plom_normal_pair64(ns, &za19, &zb19);
        const double d3_7 = sqdt * za19;
        const double w0_7 = sqdt * zb19;
        const double f7 = X[7] + ( - (r0_7*X[7]) + (r3_7))*dtr + - plom_sqrt_rate((r0_7*X[7]))*d0_7+ plom_sqrt_rate((r3_7))*d3_7 + - (r0_7*X[7])*P47*w0_7;
        const double f17 = X[17] + ( - (r1_7*X[17]) + (r0_7*X[7]))*dtr + - plom_sqrt_rate((r1_7*X[17]))*d1_7+ plom_sqrt_rate((r0_7*X[7]))*d0_7 + + (r0_7*X[7])*P47*w0_7;
        const double f27 = X[27] + ( - (r2_7*X[27]) + (r1_7*X[17]))*dtr + - plom_sqrt_rate((r2_7*X[27]))*d2_7+ plom_sqrt_rate((r1_7*X[17]))*d1_7;
        s37 += ( + (r1_7*X[17]))*dtr + + plom_sqrt_rate((r1_7*X[17]))*d1_7;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        X[17] = (f17 < 0.0) ? 0.0 : f17;
        const double r0_8 = ((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])) / C8);
        const double r1_8 = P18;
        const double r2_8 = P28;
        const double r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        double za20, zb20;
        plom_normal_pair64(ns, &za20, &zb20);
        const double d0_8 = sqdt * za20;
        const double d1_8 = sqdt * zb20;
        double za21, zb21;
        plom_normal_pair64(ns, &za21, &zb21);
        const double d2_8 = sqdt * za21;
        const double d3_8 = sqdt * zb21;
        double za22, zb22;
        plom_normal_pair64(ns, &za22, &zb22);
        const double w0_8 = sqdt * za22;
        const double f8 = X[8] + ( - (r0_8*X[8]) + (r3_8))*dtr + - plom_sqrt_rate((r0_8*X[8]))*d0_8+ plom_sqrt_rate((r3_8))*d3_8 + - (r0_8*X[8])*P48*w0_8;
        const double f18 = X[18] + ( - (r1_8*X[18]) + (r0_8*X[8]))*dtr + - plom_sqrt_rate((r1_8*X[18]))*d1_8+ plom_sqrt_rate((r0_8*X[8]))*d0_8 + + (r0_8*X[8])*P48*w0_8;
        const double f28 = X[28] + ( - (r2_8*X[28]) + (r1_8*X[18]))*dtr + - plom_sqrt_rate((r2_8*X[28]))*d2_8+ plom_sqrt_rate((r1_8*X[18]))*d1_8;
        s38 += ( + (r1_8*X[18]))*dtr + + plom_sqrt_rate((r1_8*X[18]))*d1_8;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        X[18] = (f18 < 0.0) ? 0.0 : f18;
        const double r0_9 = ((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])) / C9);
        const double r1_9 = P19;
        const double r2_9 = P29;
        const double r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        const double d0_9 = sqdt * zb22;
        double za23, zb23;
        plom_normal_pair64(ns, &za23, &zb23);
        const double d1_9 = sqdt * za23;
        const double d2_9 = sqdt * zb23;
        double za24, zb24;
        plom_normal_pair64(ns, &za24, &zb24);
        const double d3_9 = sqdt * za24;
        const double w0_9 = sqdt * zb24;
        const double f9 = X[9] + ( - (r0_9*X[9]) + (r3_9))*dtr + - plom_sqrt_rate((r0_9*X[9]))*d0_9+ plom_sqrt_rate((r3_9))*d3_9 + - (r0_9*X[9])*P49*w0_9;
        const double f19 = X[19] + ( - (r1_9*X[19]) + (r0_9*X[9]))*dtr + - plom_sqrt_rate((r1_9*X[19]))*d1_9+ plom_sqrt_rate((r0_9*X[9]))*d0_9 + + (r0_9*X[9])*P49*w0_9;
        const double f29 = X[29] + ( - (r2_9*X[29]) + (r1_9*X[19]))*dtr + - plom_sqrt_rate((r2_9*X[29]))*d2_9+ plom_sqrt_rate((r1_9*X[19]))*d1_9;
        s39 += ( + (r1_9*X[19]))*dtr + + plom_sqrt_rate((r1_9*X[19]))*d1_9;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        X[19] = (f19 < 0.0) ? 0.0 : f19;
        X[20] = (f20 < 0.0) ? 0.0 : f20;
        X[21] = (f21 < 0.0) ? 0.0 : f21;
        X[22] = (f22 < 0.0) ? 0.0 : f22;
        X[23] = (f23 < 0.0) ? 0.0 : f23;
        X[24] = (f24 < 0.0) ? 0.0 : f24;
        X[25] = (f25 < 0.0) ? 0.0 : f25;
        X[26] = (f26 < 0.0) ? 0.0 : f26;
        X[27] = (f27 < 0.0) ? 0.0 : f27;
        X[28] = (f28 < 0.0) ? 0.0 : f28;
        X[29] = (f29 < 0.0) ? 0.0 : f29;
        const double f30 = X[30] + s30;
        X[30] = (f30 < 0.0) ? 0.0 : f30;
        const double f31 = X[31] + s31;
        X[31] = (f31 < 0.0) ? 0.0 : f31;
        const double f32 = X[32] + s32;
        X[32] = (f32 < 0.0) ? 0.0 : f32;
        const double f33 = X[33] + s33;
        X[33] = (f33 < 0.0) ? 0.0 : f33;
        const double f34 = X[34] + s34;
        X[34] = (f34 < 0.0) ? 0.0 : f34;
        const double f35 = X[35] + s35;
        X[35] = (f35 < 0.0) ? 0.0 : f35;
        const double f36 = X[36] + s36;
        X[36] = (f36 < 0.0) ? 0.0 : f36;
        const double f37 = X[37] + s37;
        X[37] = (f37 < 0.0) ? 0.0 : f37;
        const double f38 = X[38] + s38;
        X[38] = (f38 < 0.0) ? 0.0 : f38;
        const double f39 = X[39] + s39;
        X[39] = (f39 < 0.0) ? 0.0 : f39;
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_interval_no_dem_sto_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
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
    const double P20 = lp[20];
    const double P21 = lp[21];
    const double P22 = lp[22];
    const double P23 = lp[23];
    const double P24 = lp[24];
    const double P25 = lp[25];
    const double P26 = lp[26];
    const double P27 = lp[27];
    const double P28 = lp[28];
    const double P29 = lp[29];
    const double P30 = lp[30];
    const double P31 = lp[31];
    const double P32 = lp[32];
    const double P33 = lp[33];
    const double P34 = lp[34];
    const double P35 = lp[35];
    const double P36 = lp[36];
    const double P37 = lp[37];
    const double P38 = lp[38];
    const double P39 = lp[39];
    const double P40 = lp[40];
    const double P41 = lp[41];
    const double P42 = lp[42];
    const double P43 = lp[43];
    const double P44 = lp[44];
    const double P45 = lp[45];
    const double P46 = lp[46];
    const double P47 = lp[47];
    const double P48 = lp[48];
    const double P49 = lp[49];
    const double P80 = PLOM_UNIFORM64(pv[80]);
    const double P81 = PLOM_UNIFORM64(pv[81]);
    const double P82 = PLOM_UNIFORM64(pv[82]);
    const double P83 = PLOM_UNIFORM64(pv[83]);
    const double P84 = PLOM_UNIFORM64(pv[84]);
    const double P85 = PLOM_UNIFORM64(pv[85]);
    const double P86 = PLOM_UNIFORM64(pv[86]);
    const double P87 = PLOM_UNIFORM64(pv[87]);
    const double P88 = PLOM_UNIFORM64(pv[88]);
    const double P89 = PLOM_UNIFORM64(pv[89]);
    const double P90 = PLOM_UNIFORM64(pv[90]);
    const double P91 = PLOM_UNIFORM64(pv[91]);
    const double P92 = PLOM_UNIFORM64(pv[92]);
    const double P93 = PLOM_UNIFORM64(pv[93]);
    const double P94 = PLOM_UNIFORM64(pv[94]);
    const double P95 = PLOM_UNIFORM64(pv[95]);
    const double P96 = PLOM_UNIFORM64(pv[96]);
    const double P97 = PLOM_UNIFORM64(pv[97]);
    const double P98 = PLOM_UNIFORM64(pv[98]);
    const double P99 = PLOM_UNIFORM64(pv[99]);
    const double P100 = PLOM_UNIFORM64(pv[100]);
    const double P101 = PLOM_UNIFORM64(pv[101]);
    const double P102 = PLOM_UNIFORM64(pv[102]);
    const double P103 = PLOM_UNIFORM64(pv[103]);
    const double P104 = PLOM_UNIFORM64(pv[104]);
    const double P105 = PLOM_UNIFORM64(pv[105]);
    const double P106 = PLOM_UNIFORM64(pv[106]);
    const double P107 = PLOM_UNIFORM64(pv[107]);
    const double P108 = PLOM_UNIFORM64(pv[108]);
    const double P109 = PLOM_UNIFORM64(pv[109]);
    const double P110 = PLOM_UNIFORM64(pv[110]);
    const double P111 = PLOM_UNIFORM64(pv[111]);
    const double P112 = PLOM_UNIFORM64(pv[112]);
    const double P113 = PLOM_UNIFORM64(pv[113]);
    const double P114 = PLOM_UNIFORM64(pv[114]);
    const double P115 = PLOM_UNIFORM64(pv[115]);
    const double P116 = PLOM_UNIFORM64(pv[116]);
    const double P117 = PLOM_UNIFORM64(pv[117]);
    const double P118 = PLOM_UNIFORM64(pv[118]);
    const double P119 = PLOM_UNIFORM64(pv[119]);
    const double P120 = PLOM_UNIFORM64(pv[120]);
    const double P121 = PLOM_UNIFORM64(pv[121]);
    const double P122 = PLOM_UNIFORM64(pv[122]);
    const double P123 = PLOM_UNIFORM64(pv[123]);
    const double P124 = PLOM_UNIFORM64(pv[124]);
    const double P125 = PLOM_UNIFORM64(pv[125]);
    const double P126 = PLOM_UNIFORM64(pv[126]);
    const double P127 = PLOM_UNIFORM64(pv[127]);
    const double P128 = PLOM_UNIFORM64(pv[128]);
    const double P129 = PLOM_UNIFORM64(pv[129]);
    const double P130 = PLOM_UNIFORM64(pv[130]);
    const double P131 = PLOM_UNIFORM64(pv[131]);
    const double P132 = PLOM_UNIFORM64(pv[132]);
    const double P133 = PLOM_UNIFORM64(pv[133]);
    const double P134 = PLOM_UNIFORM64(pv[134]);
    const double P135 = PLOM_UNIFORM64(pv[135]);
    const double P136 = PLOM_UNIFORM64(pv[136]);
    const double P137 = PLOM_UNIFORM64(pv[137]);
    const double P138 = PLOM_UNIFORM64(pv[138]);
    const double P139 = PLOM_UNIFORM64(pv[139]);
    const double P140 = PLOM_UNIFORM64(pv[140]);
    const double P141 = PLOM_UNIFORM64(pv[141]);
    const double P142 = PLOM_UNIFORM64(pv[142]);
    const double P143 = PLOM_UNIFORM64(pv[143]);
    const double P144 = PLOM_UNIFORM64(pv[144]);
    const double P145 = PLOM_UNIFORM64(pv[145]);
    const double P146 = PLOM_UNIFORM64(pv[146]);
    const double P147 = PLOM_UNIFORM64(pv[147]);
    const double P148 = PLOM_UNIFORM64(pv[148]);
    const double P149 = PLOM_UNIFORM64(pv[149]);
    const double P150 = PLOM_UNIFORM64(pv[150]);
    const double P151 = PLOM_UNIFORM64(pv[151]);
    const double P152 = PLOM_UNIFORM64(pv[152]);
    const double P153 = PLOM_UNIFORM64(pv[153]);
    const double P154 = PLOM_UNIFORM64(pv[154]);
    const double P155 = PLOM_UNIFORM64(pv[155]);
    const double P156 = PLOM_UNIFORM64(pv[156]);
    const double P157 = PLOM_UNIFORM64(pv[157]);
    const double P158 = PLOM_UNIFORM64(pv[158]);
    const double P159 = PLOM_UNIFORM64(pv[159]);
    const double P160 = PLOM_UNIFORM64(pv[160]);
    const double P161 = PLOM_UNIFORM64(pv[161]);
    const double P162 = PLOM_UNIFORM64(pv[162]);
    const double P163 = PLOM_UNIFORM64(pv[163]);
    const double P164 = PLOM_UNIFORM64(pv[164]);
    const double P165 = PLOM_UNIFORM64(pv[165]);
    const double P166 = PLOM_UNIFORM64(pv[166]);
    const double P167 = PLOM_UNIFORM64(pv[167]);
    const double P168 = PLOM_UNIFORM64(pv[168]);
    const double P169 = PLOM_UNIFORM64(pv[169]);
    const double P170 = PLOM_UNIFORM64(pv[170]);
    const double P171 = PLOM_UNIFORM64(pv[171]);
    const double P172 = PLOM_UNIFORM64(pv[172]);
    const double P173 = PLOM_UNIFORM64(pv[173]);
    const double P174 = PLOM_UNIFORM64(pv[174]);
    const double P175 = PLOM_UNIFORM64(pv[175]);
    const double P176 = PLOM_UNIFORM64(pv[176]);
    const double P177 = PLOM_UNIFORM64(pv[177]);
    const double P178 = PLOM_UNIFORM64(pv[178]);
    const double P179 = PLOM_UNIFORM64(pv[179]);
    X[30] = 0.0; /* reset_inc */
    X[31] = 0.0; /* reset_inc */
    X[32] = 0.0; /* reset_inc */
    X[33] = 0.0; /* reset_inc */
    X[34] = 0.0; /* reset_inc */
    X[35] = 0.0; /* reset_inc */
    X[36] = 0.0; /* reset_inc */
    X[37] = 0.0; /* reset_inc */
    X[38] = 0.0; /* reset_inc */
    X[39] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double C4 = plom_cov_ld(cvk + 4);
        const double C5 = plom_cov_ld(cvk + 5);
        const double C6 = plom_cov_ld(cvk + 6);
        const double C7 = plom_cov_ld(cvk + 7);
        const double C8 = plom_cov_ld(cvk + 8);
        const double C9 = plom_cov_ld(cvk + 9);
        double s30 = 0.0;
        double s31 = 0.0;
        double s32 = 0.0;
        double s33 = 0.0;
        double s34 = 0.0;
        double s35 = 0.0;
        double s36 = 0.0;
        double s37 = 0.0;
        double s38 = 0.0;
        double s39 = 0.0;
        const double r0_0 = ((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])) / C0);
        const double r1_0 = P10;
        const double r2_0 = P20;
        const double r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        double za0, zb0;
        plom_normal_pair64(ns, &za0, &zb0);
        const double w0_0 = sqdt * za0;
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r3_0))*dtr + - (r0_0*X[0])*P40*w0_0;
        const double f10 = X[10] + ( - (r1_0*X[10]) + (r0_0*X[0]))*dtr + + (r0_0*X[0])*P40*w0_0;
        const double f20 = X[20] + ( - (r2_0*X[20]) + (r1_0*X[10]))*dtr;
        s30 += ( + (r1_0*X[10]))*dtr;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        const double r0_1 = ((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])) / C1);
        const double r1_1 = P11;
        const double r2_1 = P21;
        const double r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        const double w0_1 = sqdt * zb0;
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r3_1))*dtr + - (r0_1*X[1])*P41*w0_1;
        const double f11 = X[11] + ( - (r1_1*X[11]) + (r0_1*X[1]))*dtr + + (r0_1*X[1])*P41*w0_1;
        const double f21 = X[21] + ( - (r2_1*X[21]) + (r1_1*X[11]))*dtr;
        s31 += ( + (r1_1*X[11]))*dtr;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        const double r0_2 = ((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])) / C2);
        const double r1_2 = P12;
        const double r2_2 = P22;
        const double r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        double za1, zb1;
        plom_normal_pair64(ns, &za1, &zb1);
        const double w0_2 = sqdt * za1;
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r3_2))*dtr + - (r0_2*X[2])*P42*w0_2;
        const double f12 = X[12] + ( - (r1_2*X[12]) + (r0_2*X[2]))*dtr + + (r0_2*X[2])*P42*w0_2;
        const double f22 = X[22] + ( - (r2_2*X[22]) + (r1_2*X[12]))*dtr;
        s32 += ( + (r1_2*X[12]))*dtr;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        X[12] = (f12 < 0.0) ? 0.0 : f12;
        const double r0_3 = ((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])) / C3);
        const double r1_3 = P13;
        const double r2_3 = P23;
        const double r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        const double w0_3 = sqdt * zb1;
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r3_3))*dtr + - (r0_3*X[3])*P43*w0_3;
        const double f13 = X[13] + ( - (r1_3*X[13]) + (r0_3*X[3]))*dtr + + (r0_3*X[3])*P43*w0_3;
        const double f23 = X[23] + ( - (r2_3*X[23]) + (r1_3*X[13]))*dtr;
        s33 += ( + (r1_3*X[13]))*dtr;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[13] = (f13 < 0.0) ? 0.0 : f13;
        const double r0_4 = ((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])) / C4);
        const double r1_4 = P14;
        const double r2_4 = P24;
        const double r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        double za2, zb2;
        plom_normal_pair64(ns, &za2, &zb2);
        const double w0_4 = sqdt * za2;
        const double f4 = X[4] + ( - (r0_4*X[4]) + (r3_4))*dtr + - (r0_4*X[4])*P44*w0_4;
        const double f14 = X[14] + ( - (r1_4*X[14]) + (r0_4*X[4]))*dtr + + (r0_4*X[4])*P44*w0_4;
        const double f24 = X[24] + ( - (r2_4*X[24]) + (r1_4*X[14]))*dtr;
        s34 += ( + (r1_4*X[14]))*dtr;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[14] = (f14 < 0.0) ? 0.0 : f14;
        const double r0_5 = ((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])) / C5);
        const double r1_5 = P15;
        const double r2_5 = P25;
        const double r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        const double w0_5 = sqdt * zb2;
        const double f5 = X[5] + ( - (r0_5*X[5]) + (r3_5))*dtr + - (r0_5*X[5])*P45*w0_5;
        const double f15 = X[15] + ( - (r1_5*X[15]) + (r0_5*X[5]))*dtr + + (r0_5*X[5])*P45*w0_5;
        const double f25 = X[25] + ( - (r2_5*X[25]) + (r1_5*X[15]))*dtr;
        s35 += ( + (r1_5*X[15]))*dtr;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[15] = (f15 < 0.0) ? 0.0 : f15;
        const double r0_6 = ((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])) / C6);
        const double r1_6 = P16;
        const double r2_6 = P26;
        const double r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        double za3, zb3;
        plom_normal_pair64(ns, &za3, &zb3);
        const double w0_6 = sqdt * za3;
        const double f6 = X[6] + ( - (r0_6*X[6]) + (r3_6))*dtr + - (r0_6*X[6])*P46*w0_6;
        const double f16 = X[16] + ( - (r1_6*X[16]) + (r0_6*X[6]))*dtr + + (r0_6*X[6])*P46*w0_6;
        const double f26 = X[26] + ( - (r2_6*X[26]) + (r1_6*X[16]))*dtr;
        s36 += ( + (r1_6*X[16]))*dtr;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[16] = (f16 < 0.0) ? 0.0 : f16;
        const double r0_7 = ((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])) / C7);
        const double r1_7 = P17;
        const double r2_7 = P27;
        const double r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        const double w0_7 = sqdt * zb3;
        const double f7 = X[7] + ( - (r0_7*X[7]) + (r3_7))*dtr + - (r0_7*X[7])*P47*w0_7;
        const double f17 = X[17] + ( - (r1_7*X[17]) + (r0_7*X[7]))*dtr + + (r0_7*X[7])*P47*w0_7;
        const double f27 = X[27] + ( - (r2_7*X[27]) + (r1_7*X[17]))*dtr;
        s37 += ( + (r1_7*X[17]))*dtr;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        X[17] = (f17 < 0.0) ? 0.0 : f17;
        const double r0_8 = ((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])) / C8);
        const double r1_8 = P18;
        const double r2_8 = P28;
        const double r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        double za4, zb4;
        plom_normal_pair64(ns, &za4, &zb4);
        const double w0_8 = sqdt * za4;
        const double f8 = X[8] + ( - (r0_8*X[8]) + (r3_8))*dtr + - (r0_8*X[8])*P48*w0_8;
        const double f18 = X[18] + ( - (r1_8*X[18]) + (r0_8*X[8]))*dtr + + (r0_8*X[8])*P48*w0_8;
        const double f28 = X[28] + ( - (r2_8*X[28]) + (r1_8*X[18]))*dtr;
        s38 += ( + (r1_8*X[18]))*dtr;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        X[18] = (f18 < 0.0) ? 0.0 : f18;
        const double r0_9 = ((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])) / C9);
        const double r1_9 = P19;
        const double r2_9 = P29;
        const double r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        const double w0_9 = sqdt * zb4;
        const double f9 = X[9] + ( - (r0_9*X[9]) + (r3_9))*dtr + - (r0_9*X[9])*P49*w0_9;
        const double f19 = X[19] + ( - (r1_9*X[19]) + (r0_9*X[9]))*dtr + + (r0_9*X[9])*P49*w0_9;
        const double f29 = X[29] + ( - (r2_9*X[29]) + (r1_9*X[19]))*dtr;
        s39 += ( + (r1_9*X[19]))*dtr;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        X[19] = (f19 < 0.0) ? 0.0 : f19;
        X[20] = (f20 < 0.0) ? 0.0 : f20;
        X[21] = (f21 < 0.0) ? 0.0 : f21;
        X[22] = (f22 < 0.0) ? 0.0 : f22;
        X[23] = (f23 < 0.0) ? 0.0 : f23;
        X[24] = (f24 < 0.0) ? 0.0 : f24;
        X[25] = (f25 < 0.0) ? 0.0 : f25;
        X[26] = (f26 < 0.0) ? 0.0 : f26;
        X[27] = (f27 < 0.0) ? 0.0 : f27;
        X[28] = (f28 < 0.0) ? 0.0 : f28;
        X[29] = (f29 < 0.0) ? 0.0 : f29;
        const double f30 = X[30] + s30;
        X[30] = (f30 < 0.0) ? 0.0 : f30;
        const double f31 = X[31] + s31;
        X[31] = (f31 < 0.0) ? 0.0 : f31;
        const double f32 = X[32] + s32;
        X[32] = (f32 < 0.0) ? 0.0 : f32;
        const double f33 = X[33] + s33;
        X[33] = (f33 < 0.0) ? 0.0 : f33;
        const double f34 = X[34] + s34;
        X[34] = (f34 < 0.0) ? 0.0 : f34;
        const double f35 = X[35] + s35;
        X[35] = (f35 < 0.0) ? 0.0 : f35;
        const double f36 = X[36] + s36;
        X[36] = (f36 < 0.0) ? 0.0 : f36;
        const double f37 = X[37] + s37;
        X[37] = (f37 < 0.0) ? 0.0 : f37;
        const double f38 = X[38] + s38;
        X[38] = (f38 < 0.0) ? 0.0 : f38;
        const double f39 = X[39] + s39;
        X[39] = (f39 < 0.0) ? 0.0 : f39;
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_interval_no_env_sto_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
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
    const double P20 = lp[20];
    const double P21 = lp[21];
    const double P22 = lp[22];
    const double P23 = lp[23];
    const double P24 = lp[24];
    const double P25 = lp[25];
    const double P26 = lp[26];
    const double P27 = lp[27];
    const double P28 = lp[28];
    const double P29 = lp[29];
    const double P30 = lp[30];
    const double P31 = lp[31];
    const double P32 = lp[32];
    const double P33 = lp[33];
    const double P34 = lp[34];
    const double P35 = lp[35];
    const double P36 = lp[36];
    const double P37 = lp[37];
    const double P38 = lp[38];
    const double P39 = lp[39];
    const double P80 = PLOM_UNIFORM64(pv[80]);
    const double P81 = PLOM_UNIFORM64(pv[81]);
    const double P82 = PLOM_UNIFORM64(pv[82]);
    const double P83 = PLOM_UNIFORM64(pv[83]);
    const double P84 = PLOM_UNIFORM64(pv[84]);
    const double P85 = PLOM_UNIFORM64(pv[85]);
    const double P86 = PLOM_UNIFORM64(pv[86]);
    const double P87 = PLOM_UNIFORM64(pv[87]);
    const double P88 = PLOM_UNIFORM64(pv[88]);
    const double P89 = PLOM_UNIFORM64(pv[89]);
    const double P90 = PLOM_UNIFORM64(pv[90]);
    const double P91 = PLOM_UNIFORM64(pv[91]);
    const double P92 = PLOM_UNIFORM64(pv[92]);
    const double P93 = PLOM_UNIFORM64(pv[93]);
    const double P94 = PLOM_UNIFORM64(pv[94]);
    const double P95 = PLOM_UNIFORM64(pv[95]);
    const double P96 = PLOM_UNIFORM64(pv[96]);
    const double P97 = PLOM_UNIFORM64(pv[97]);
    const double P98 = PLOM_UNIFORM64(pv[98]);
    const double P99 = PLOM_UNIFORM64(pv[99]);
    const double P100 = PLOM_UNIFORM64(pv[100]);
    const double P101 = PLOM_UNIFORM64(pv[101]);
    const double P102 = PLOM_UNIFORM64(pv[102]);
    const double P103 = PLOM_UNIFORM64(pv[103]);
    const double P104 = PLOM_UNIFORM64(pv[104]);
    const double P105 = PLOM_UNIFORM64(pv[105]);
    const double P106 = PLOM_UNIFORM64(pv[106]);
    const double P107 = PLOM_UNIFORM64(pv[107]);
    const double P108 = PLOM_UNIFORM64(pv[108]);
    const double P109 = PLOM_UNIFORM64(pv[109]);
    const double P110 = PLOM_UNIFORM64(pv[110]);
    const double P111 = PLOM_UNIFORM64(pv[111]);
    const double P112 = PLOM_UNIFORM64(pv[112]);
    const double P113 = PLOM_UNIFORM64(pv[113]);
    const double P114 = PLOM_UNIFORM64(pv[114]);
    const double P115 = PLOM_UNIFORM64(pv[115]);
    const double P116 = PLOM_UNIFORM64(pv[116]);
    const double P117 = PLOM_UNIFORM64(pv[117]);
    const double P118 = PLOM_UNIFORM64(pv[118]);
    const double P119 = PLOM_UNIFORM64(pv[119]);
    const double P120 = PLOM_UNIFORM64(pv[120]);
    const double P121 = PLOM_UNIFORM64(pv[121]);
    const double P122 = PLOM_UNIFORM64(pv[122]);
    const double P123 = PLOM_UNIFORM64(pv[123]);
    const double P124 = PLOM_UNIFORM64(pv[124]);
    const double P125 = PLOM_UNIFORM64(pv[125]);
    const double P126 = PLOM_UNIFORM64(pv[126]);
    const double P127 = PLOM_UNIFORM64(pv[127]);
    const double P128 = PLOM_UNIFORM64(pv[128]);
    const double P129 = PLOM_UNIFORM64(pv[129]);
    const double P130 = PLOM_UNIFORM64(pv[130]);
    const double P131 = PLOM_UNIFORM64(pv[131]);
    const double P132 = PLOM_UNIFORM64(pv[132]);
    const double P133 = PLOM_UNIFORM64(pv[133]);
    const double P134 = PLOM_UNIFORM64(pv[134]);
    const double P135 = PLOM_UNIFORM64(pv[135]);
    const double P136 = PLOM_UNIFORM64(pv[136]);
    const double P137 = PLOM_UNIFORM64(pv[137]);
    const double P138 = PLOM_UNIFORM64(pv[138]);
    const double P139 = PLOM_UNIFORM64(pv[139]);
    const double P140 = PLOM_UNIFORM64(pv[140]);
    const double P141 = PLOM_UNIFORM64(pv[141]);
    const double P142 = PLOM_UNIFORM64(pv[142]);
    const double P143 = PLOM_UNIFORM64(pv[143]);
    const double P144 = PLOM_UNIFORM64(pv[144]);
    const double P145 = PLOM_UNIFORM64(pv[145]);
    const double P146 = PLOM_UNIFORM64(pv[146]);
    const double P147 = PLOM_UNIFORM64(pv[147]);
    const double P148 = PLOM_UNIFORM64(pv[148]);
    const double P149 = PLOM_UNIFORM64(pv[149]);
    const double P150 = PLOM_UNIFORM64(pv[150]);
    const double P151 = PLOM_UNIFORM64(pv[151]);
    const double P152 = PLOM_UNIFORM64(pv[152]);
    const double P153 = PLOM_UNIFORM64(pv[153]);
    const double P154 = PLOM_UNIFORM64(pv[154]);
    const double P155 = PLOM_UNIFORM64(pv[155]);
    const double P156 = PLOM_UNIFORM64(pv[156]);
    const double P157 = PLOM_UNIFORM64(pv[157]);
    const double P158 = PLOM_UNIFORM64(pv[158]);
    const double P159 = PLOM_UNIFORM64(pv[159]);
    const double P160 = PLOM_UNIFORM64(pv[160]);
    const double P161 = PLOM_UNIFORM64(pv[161]);
    const double P162 = PLOM_UNIFORM64(pv[162]);
    const double P163 = PLOM_UNIFORM64(pv[163]);
    const double P164 = PLOM_UNIFORM64(pv[164]);
    const double P165 = PLOM_UNIFORM64(pv[165]);
    const double P166 = PLOM_UNIFORM64(pv[166]);
    const double P167 = PLOM_UNIFORM64(pv[167]);
    const double P168 = PLOM_UNIFORM64(pv[168]);
    const double P169 = PLOM_UNIFORM64(pv[169]);
    const double P170 = PLOM_UNIFORM64(pv[170]);
    const double P171 = PLOM_UNIFORM64(pv[171]);
    const double P172 = PLOM_UNIFORM64(pv[172]);
    const double P173 = PLOM_UNIFORM64(pv[173]);
    const double P174 = PLOM_UNIFORM64(pv[174]);
    const double P175 = PLOM_UNIFORM64(pv[175]);
    const double P176 = PLOM_UNIFORM64(pv[176]);
    const double P177 = PLOM_UNIFORM64(pv[177]);
    const double P178 = PLOM_UNIFORM64(pv[178]);
    const double P179 = PLOM_UNIFORM64(pv[179]);
    X[30] = 0.0; /* reset_inc */
    X[31] = 0.0; /* reset_inc */
    X[32] = 0.0; /* reset_inc */
    X[33] = 0.0; /* reset_inc */
    X[34] = 0.0; /* reset_inc */
    X[35] = 0.0; /* reset_inc */
    X[36] = 0.0; /* reset_inc */
    X[37] = 0.0; /* reset_inc */
    X[38] = 0.0; /* reset_inc */
    X[39] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double C4 = plom_cov_ld(cvk + 4);
        const double C5 = plom_cov_ld(cvk + 5);
        const double C6 = plom_cov_ld(cvk + 6);
        const double C7 = plom_cov_ld(cvk + 7);
        const double C8 = plom_cov_ld(cvk + 8);
        const double C9 = plom_cov_ld(cvk + 9);
        double s30 = 0.0;
        double s31 = 0.0;
        double s32 = 0.0;
        double s33 = 0.0;
        double s34 = 0.0;
        double s35 = 0.0;
        double s36 = 0.0;
        double s37 = 0.0;
        double s38 = 0.0;
        double s39 = 0.0;
        const double r0_0 = ((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])) / C0);
        const double r1_0 = P10;
        const double r2_0 = P20;
        const double r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        double za0, zb0;
        plom_normal_pair64(ns, &za0, &zb0);
        const double d0_0 = sqdt * za0;
        const double d1_0 = sqdt * zb0;
        double za1, zb1;
        plom_normal_pair64(ns, &za1, &zb1);
        const double d2_0 = sqdt * za1;
        const double d3_0 = sqdt * zb1;
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r3_0))*dtr + - plom_sqrt_rate((r0_0*X[0]))*d0_0+ plom_sqrt_rate((r3_0))*d3_0;
        const double f10 = X[10] + ( - (r1_0*X[10]) + (r0_0*X[0]))*dtr + - plom_sqrt_rate((r1_0*X[10]))*d1_0+ plom_sqrt_rate((r0_0*X[0]))*d0_0;
        const double f20 = X[20] + ( - (r2_0*X[20]) + (r1_0*X[10]))*dtr + - plom_sqrt_rate((r2_0*X[20]))*d2_0+ plom_sqrt_rate((r1_0*X[10]))*d1_0;
        s30 += ( + (r1_0*X[10]))*dtr + + plom_sqrt_rate((r1_0*X[10]))*d1_0;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        const double r0_1 = ((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])) / C1);
        const double r1_1 = P11;
        const double r2_1 = P21;
        const double r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        double za2, zb2;
        plom_normal_pair64(ns, &za2, &zb2);
        const double d0_1 = sqdt * za2;
        const double d1_1 = sqdt * zb2;
        double za3, zb3;
        plom_normal_pair64(ns, &za3, &zb3);
        const double d2_1 = sqdt * za3;
        const double d3_1 = sqdt * zb3;
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r3_1))*dtr + - plom_sqrt_rate((r0_1*X[1]))*d0_1+ plom_sqrt_rate((r3_1))*d3_1;
        const double f11 = X[11] + ( - (r1_1*X[11]) + (r0_1*X[1]))*dtr + - plom_sqrt_rate((r1_1*X[11]))*d1_1+ plom_sqrt_rate((r0_1*X[1]))*d0_1;
        const double f21 = X[21] + ( - (r2_1*X[21]) + (r1_1*X[11]))*dtr + - plom_sqrt_rate((r2_1*X[21]))*d2_1+ plom_sqrt_rate((r1_1*X[11]))*d1_1;
        s31 += ( + (r1_1*X[11]))*dtr + + plom_sqrt_rate((r1_1*X[11]))*d1_1;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        const double r0_2 = ((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])) / C2);
        const double r1_2 = P12;
        const double r2_2 = P22;
        const double r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        double za4, zb4;
        plom_normal_pair64(ns, &za4, &zb4);
        const double d0_2 = sqdt * za4;
        const double d1_2 = sqdt * zb4;
        double za5, zb5;
        plom_normal_pair64(ns, &za5, &zb5);
        const double d2_2 = sqdt * za5;
        const double d3_2 = sqdt * zb5;
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r3_2))*dtr + - plom_sqrt_rate((r0_2*X[2]))*d0_2+ plom_sqrt_rate((r3_2))*d3_2;
        const double f12 = X[12] + ( - (r1_2*X[12]) + (r0_2*X[2]))*dtr + - plom_sqrt_rate((r1_2*X[12]))*d1_2+ plom_sqrt_rate((r0_2*X[2]))*d0_2;
        const double f22 = X[22] + ( - (r2_2*X[22]) + (r1_2*X[12]))*dtr + - plom_sqrt_rate((r2_2*X[22]))*d2_2+ plom_sqrt_rate((r1_2*X[12]))*d1_2;
        s32 += ( + (r1_2*X[12]))*dtr + + plom_sqrt_rate((r1_2*X[12]))*d1_2;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        X[12] = (f12 < 0.0) ? 0.0 : f12;
        const double r0_3 = ((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])) / C3);
        const double r1_3 = P13;
        const double r2_3 = P23;
        const double r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        double za6, zb6;
        plom_normal_pair64(ns, &za6, &zb6);
        const double d0_3 = sqdt * za6;
        const double d1_3 = sqdt * zb6;
        double za7, zb7;
        plom_normal_pair64(ns, &za7, &zb7);
        const double d2_3 = sqdt * za7;
        const double d3_3 = sqdt * zb7;
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r3_3))*dtr + - plom_sqrt_rate((r0_3*X[3]))*d0_3+ plom_sqrt_rate((r3_3))*d3_3;
        const double f13 = X[13] + ( - (r1_3*X[13]) + (r0_3*X[3]))*dtr + - plom_sqrt_rate((r1_3*X[13]))*d1_3+ plom_sqrt_rate((r0_3*X[3]))*d0_3;
        const double f23 = X[23] + ( - (r2_3*X[23]) + (r1_3*X[13]))*dtr + - plom_sqrt_rate((r2_3*X[23]))*d2_3+ plom_sqrt_rate((r1_3*X[13]))*d1_3;
        s33 += ( + (r1_3*X[13]))*dtr + + plom_sqrt_rate((r1_3*X[13]))*d1_3;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[13] = (f13 < 0.0) ? 0.0 : f13;
        const double r0_4 = ((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])) / C4);
        const double r1_4 = P14;
        const double r2_4 = P24;
        const double r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        double za8, zb8;
        plom_normal_pair64(ns, &za8, &zb8);
        const double d0_4 = sqdt * za8;
        const double d1_4 = sqdt * zb8;
        double za9, zb9;
        plom_normal_pair64(ns, &za9, &zb9);
        const double d2_4 = sqdt * za9;
        const double d3_4 = sqdt * zb9;
        const double f4 = X[4] + ( - (r0_4*X[4]) + (r3_4))*dtr + - plom_sqrt_rate((r0_4*X[4]))*d0_4+ plom_sqrt_rate((r3_4))*d3_4;
        const double f14 = X[14] + ( - (r1_4*X[14]) + (r0_4*X[4]))*dtr + - plom_sqrt_rate((r1_4*X[14]))*d1_4+ plom_sqrt_rate((r0_4*X[4]))*d0_4;
        const double f24 = X[24] + ( - (r2_4*X[24]) + (r1_4*X[14]))*dtr + - plom_sqrt_rate((r2_4*X[24]))*d2_4+ plom_sqrt_rate((r1_4*X[14]))*d1_4;
        s34 += ( + (r1_4*X[14]))*dtr + + plom_sqrt_rate((r1_4*X[14]))*d1_4;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[14] = (f14 < 0.0) ? 0.0 : f14;
        const double r0_5 = ((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])) / C5);
        const double r1_5 = P15;
        const double r2_5 = P25;
        const double r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        double za10, zb10;
        plom_normal_pair64(ns, &za10, &zb10);
        const double d0_5 = sqdt * za10;
        const double d1_5 = sqdt * zb10;
        double za11, zb11;
        plom_normal_pair64(ns, &za11, &zb11);
        const double d2_5 = sqdt * za11;
        const double d3_5 = sqdt * zb11;
        const double f5 = X[5] + ( - (r0_5*X[5]) + (r3_5))*dtr + - plom_sqrt_rate((r0_5*X[5]))*d0_5+ plom_sqrt_rate((r3_5))*d3_5;
        const double f15 = X[15] + ( - (r1_5*X[15]) + (r0_5*X[5]))*dtr + - plom_sqrt_rate((r1_5*X[15]))*d1_5+ plom_sqrt_rate((r0_5*X[5]))*d0_5;
        const double f25 = X[25] + ( - (r2_5*X[25]) + (r1_5*X[15]))*dtr + - plom_sqrt_rate((r2_5*X[25]))*d2_5+ plom_sqrt_rate((r1_5*X[15]))*d1_5;
        s35 += ( + (r1_5*X[15]))*dtr + + plom_sqrt_rate((r1_5*X[15]))*d1_5;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[15] = (f15 < 0.0) ? 0.0 : f15;
        const double r0_6 = ((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])) / C6);
        const double r1_6 = P16;
        const double r2_6 = P26;
        const double r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        double za12, zb12;
        plom_normal_pair64(ns, &za12, &zb12);
        const double d0_6 = sqdt * za12;
        const double d1_6 = sqdt * zb12;
        double za13, zb13;
        plom_normal_pair64(ns, &za13, &zb13);
        const double d2_6 = sqdt * za13;
        const double d3_6 = sqdt * zb13;
        const double f6 = X[6] + ( - (r0_6*X[6]) + (r3_6))*dtr + - plom_sqrt_rate((r0_6*X[6]))*d0_6+ plom_sqrt_rate((r3_6))*d3_6;
        const double f16 = X[16] + ( - (r1_6*X[16]) + (r0_6*X[6]))*dtr + - plom_sqrt_rate((r1_6*X[16]))*d1_6+ plom_sqrt_rate((r0_6*X[6]))*d0_6;
        const double f26 = X[26] + ( - (r2_6*X[26]) + (r1_6*X[16]))*dtr + - plom_sqrt_rate((r2_6*X[26]))*d2_6+ plom_sqrt_rate((r1_6*X[16]))*d1_6;
        s36 += ( + (r1_6*X[16]))*dtr + + plom_sqrt_rate((r1_6*X[16]))*d1_6;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[16] = (f16 < 0.0) ? 0.0 : f16;
        const double r0_7 = ((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])) / C7);
        const double r1_7 = P17;
        const double r2_7 = P27;
        const double r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        double za14, zb14;
        plom_normal_pair64(ns, &za14, &zb14);
        const double d0_7 = sqdt * za14;
        const double d1_7 = sqdt * zb14;
        double za15, zb15;
        plom_normal_pair64(ns, &za15, &zb15);
        const double d2_7 = sqdt * za15;
        const double d3_7 = sqdt * zb15;
        const double f7 = X[7] + ( - (r0_7*X[7]) + (r3_7))*dtr + - plom_sqrt_rate((r0_7*X[7]))*d0_7+ plom_sqrt_rate((r3_7))*d3_7;
        const double f17 = X[17] + ( - (r1_7*X[17]) + (r0_7*X[7]))*dtr + - plom_sqrt_rate((r1_7*X[17]))*d1_7+ plom_sqrt_rate((r0_7*X[7]))*d0_7;
        const double f27 = X[27] + ( - (r2_7*X[27]) + (r1_7*X[17]))*dtr + - plom_sqrt_rate((r2_7*X[27]))*d2_7+ plom_sqrt_rate((r1_7*X[17]))*d1_7;
        s37 += ( + (r1_7*X[17]))*dtr + + plom_sqrt_rate((r1_7*X[17]))*d1_7;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        X[17] = (f17 < 0.0) ? 0.0 : f17;
        const double r0_8 = ((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])) / C8);
        const double r1_8 = P18;
        const double r2_8 = P28;
        const double r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        double za16, zb16;
        plom_normal_pair64(ns, &za16, &zb16);
        const double d0_8 = sqdt * za16;
        const double d1_8 = sqdt * zb16;
        double za17, zb17;
        plom_normal_pair64(ns, &za17, &zb17);
        const double d2_8 = sqdt * za17;
        const double d3_8 = sqdt * zb17;
        const double f8 = X[8] + ( - (r0_8*X[8]) + (r3_8))*dtr + - plom_sqrt_rate((r0_8*X[8]))*d0_8+ plom_sqrt_rate((r3_8))*d3_8;
        const double f18 = X[18] + ( - (r1_8*X[18]) + (r0_8*X[8]))*dtr + - plom_sqrt_rate((r1_8*X[18]))*d1_8+ plom_sqrt_rate((r0_8*X[8]))*d0_8;
        const double f28 = X[28] + ( - (r2_8*X[28]) + (r1_8*X[18]))*dtr + - plom_sqrt_rate((r2_8*X[28]))*d2_8+ plom_sqrt_rate((r1_8*X[18]))*d1_8;
        s38 += ( + (r1_8*X[18]))*dtr + + plom_sqrt_rate((r1_8*X[18]))*d1_8;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        X[18] = (f18 < 0.0) ? 0.0 : f18;
        const double r0_9 = ((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])) / C9);
        const double r1_9 = P19;
        const double r2_9 = P29;
        const double r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        double za18, zb18;
        plom_normal_pair64(ns, &za18, &zb18);
        const double d0_9 = sqdt * za18;
        const double d1_9 = sqdt * zb18;
        double za19, zb19;
        plom_normal_pair64(ns, &za19, &zb19);
        const double d2_9 = sqdt * za19;
        const double d3_9 = sqdt * zb19;
        const double f9 = X[9] + ( - (r0_9*X[9]) + (r3_9))*dtr + - plom_sqrt_rate((r0_9*X[9]))*d0_9+ plom_sqrt_rate((r3_9))*d3_9;
        const double f19 = X[19] + ( - (r1_9*X[19]) + (r0_9*X[9]))*dtr + - plom_sqrt_rate((r1_9*X[19]))*d1_9+ plom_sqrt_rate((r0_9*X[9]))*d0_9;
        const double f29 = X[29] + ( - (r2_9*X[29]) + (r1_9*X[19]))*dtr + - plom_sqrt_rate((r2_9*X[29]))*d2_9+ plom_sqrt_rate((r1_9*X[19]))*d1_9;
        s39 += ( + (r1_9*X[19]))*dtr + + plom_sqrt_rate((r1_9*X[19]))*d1_9;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        X[19] = (f19 < 0.0) ? 0.0 : f19;
        X[20] = (f20 < 0.0) ? 0.0 : f20;
        X[21] = (f21 < 0.0) ? 0.0 : f21;
        X[22] = (f22 < 0.0) ? 0.0 : f22;
        X[23] = (f23 < 0.0) ? 0.0 : f23;
        X[24] = (f24 < 0.0) ? 0.0 : f24;
        X[25] = (f25 < 0.0) ? 0.0 : f25;
        X[26] = (f26 < 0.0) ? 0.0 : f26;
        X[27] = (f27 < 0.0) ? 0.0 : f27;
        X[28] = (f28 < 0.0) ? 0.0 : f28;
        X[29] = (f29 < 0.0) ? 0.0 : f29;
        const double f30 = X[30] + s30;
        X[30] = (f30 < 0.0) ? 0.0 : f30;
        const double f31 = X[31] + s31;
        X[31] = (f31 < 0.0) ? 0.0 : f31;
        const double f32 = X[32] + s32;
        X[32] = (f32 < 0.0) ? 0.0 : f32;
        const double f33 = X[33] + s33;
        X[33] = (f33 < 0.0) ? 0.0 : f33;
        const double f34 = X[34] + s34;
        X[34] = (f34 < 0.0) ? 0.0 : f34;
        const double f35 = X[35] + s35;
        X[35] = (f35 < 0.0) ? 0.0 : f35;
        const double f36 = X[36] + s36;
        X[36] = (f36 < 0.0) ? 0.0 : f36;
        const double f37 = X[37] + s37;
        X[37] = (f37 < 0.0) ? 0.0 : f37;
        const double f38 = X[38] + s38;
        X[38] = (f38 < 0.0) ? 0.0 : f38;
        const double f39 = X[39] + s39;
        X[39] = (f39 < 0.0) ? 0.0 : f39;
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_interval_no_dem_sto_no_env_sto_pp_f64(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
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
    const double P20 = lp[20];
    const double P21 = lp[21];
    const double P22 = lp[22];
    const double P23 = lp[23];
    const double P24 = lp[24];
    const double P25 = lp[25];
    const double P26 = lp[26];
    const double P27 = lp[27];
    const double P28 = lp[28];
    const double P29 = lp[29];
    const double P30 = lp[30];
    const double P31 = lp[31];
    const double P32 = lp[32];
    const double P33 = lp[33];
    const double P34 = lp[34];
    const double P35 = lp[35];
    const double P36 = lp[36];
    const double P37 = lp[37];
    const double P38 = lp[38];
    const double P39 = lp[39];
    const double P80 = PLOM_UNIFORM64(pv[80]);
    const double P81 = PLOM_UNIFORM64(pv[81]);
    const double P82 = PLOM_UNIFORM64(pv[82]);
    const double P83 = PLOM_UNIFORM64(pv[83]);
    const double P84 = PLOM_UNIFORM64(pv[84]);
    const double P85 = PLOM_UNIFORM64(pv[85]);
    const double P86 = PLOM_UNIFORM64(pv[86]);
    const double P87 = PLOM_UNIFORM64(pv[87]);
    const double P88 = PLOM_UNIFORM64(pv[88]);
    const double P89 = PLOM_UNIFORM64(pv[89]);
    const double P90 = PLOM_UNIFORM64(pv[90]);
    const double P91 = PLOM_UNIFORM64(pv[91]);
    const double P92 = PLOM_UNIFORM64(pv[92]);
    const double P93 = PLOM_UNIFORM64(pv[93]);
    const double P94 = PLOM_UNIFORM64(pv[94]);
    const double P95 = PLOM_UNIFORM64(pv[95]);
    const double P96 = PLOM_UNIFORM64(pv[96]);
    const double P97 = PLOM_UNIFORM64(pv[97]);
    const double P98 = PLOM_UNIFORM64(pv[98]);
    const double P99 = PLOM_UNIFORM64(pv[99]);
    const double P100 = PLOM_UNIFORM64(pv[100]);
    const double P101 = PLOM_UNIFORM64(pv[101]);
    const double P102 = PLOM_UNIFORM64(pv[102]);
    const double P103 = PLOM_UNIFORM64(pv[103]);
    const double P104 = PLOM_UNIFORM64(pv[104]);
    const double P105 = PLOM_UNIFORM64(pv[105]);
    const double P106 = PLOM_UNIFORM64(pv[106]);
    const double P107 = PLOM_UNIFORM64(pv[107]);
    const double P108 = PLOM_UNIFORM64(pv[108]);
    const double P109 = PLOM_UNIFORM64(pv[109]);
    const double P110 = PLOM_UNIFORM64(pv[110]);
    const double P111 = PLOM_UNIFORM64(pv[111]);
    const double P112 = PLOM_UNIFORM64(pv[112]);
    const double P113 = PLOM_UNIFORM64(pv[113]);
    const double P114 = PLOM_UNIFORM64(pv[114]);
    const double P115 = PLOM_UNIFORM64(pv[115]);
    const double P116 = PLOM_UNIFORM64(pv[116]);
    const double P117 = PLOM_UNIFORM64(pv[117]);
    const double P118 = PLOM_UNIFORM64(pv[118]);
    const double P119 = PLOM_UNIFORM64(pv[119]);
    const double P120 = PLOM_UNIFORM64(pv[120]);
    const double P121 = PLOM_UNIFORM64(pv[121]);
    const double P122 = PLOM_UNIFORM64(pv[122]);
    const double P123 = PLOM_UNIFORM64(pv[123]);
    const double P124 = PLOM_UNIFORM64(pv[124]);
    const double P125 = PLOM_UNIFORM64(pv[125]);
    const double P126 = PLOM_UNIFORM64(pv[126]);
    const double P127 = PLOM_UNIFORM64(pv[127]);
    const double P128 = PLOM_UNIFORM64(pv[128]);
    const double P129 = PLOM_UNIFORM64(pv[129]);
    const double P130 = PLOM_UNIFORM64(pv[130]);
    const double P131 = PLOM_UNIFORM64(pv[131]);
    const double P132 = PLOM_UNIFORM64(pv[132]);
    const double P133 = PLOM_UNIFORM64(pv[133]);
    const double P134 = PLOM_UNIFORM64(pv[134]);
    const double P135 = PLOM_UNIFORM64(pv[135]);
    const double P136 = PLOM_UNIFORM64(pv[136]);
    const double P137 = PLOM_UNIFORM64(pv[137]);
    const double P138 = PLOM_UNIFORM64(pv[138]);
    const double P139 = PLOM_UNIFORM64(pv[139]);
    const double P140 = PLOM_UNIFORM64(pv[140]);
    const double P141 = PLOM_UNIFORM64(pv[141]);
    const double P142 = PLOM_UNIFORM64(pv[142]);
    const double P143 = PLOM_UNIFORM64(pv[143]);
    const double P144 = PLOM_UNIFORM64(pv[144]);
    const double P145 = PLOM_UNIFORM64(pv[145]);
    const double P146 = PLOM_UNIFORM64(pv[146]);
    const double P147 = PLOM_UNIFORM64(pv[147]);
    const double P148 = PLOM_UNIFORM64(pv[148]);
    const double P149 = PLOM_UNIFORM64(pv[149]);
    const double P150 = PLOM_UNIFORM64(pv[150]);
    const double P151 = PLOM_UNIFORM64(pv[151]);
    const double P152 = PLOM_UNIFORM64(pv[152]);
    const double P153 = PLOM_UNIFORM64(pv[153]);
    const double P154 = PLOM_UNIFORM64(pv[154]);
    const double P155 = PLOM_UNIFORM64(pv[155]);
    const double P156 = PLOM_UNIFORM64(pv[156]);
    const double P157 = PLOM_UNIFORM64(pv[157]);
    const double P158 = PLOM_UNIFORM64(pv[158]);
    const double P159 = PLOM_UNIFORM64(pv[159]);
    const double P160 = PLOM_UNIFORM64(pv[160]);
    const double P161 = PLOM_UNIFORM64(pv[161]);
    const double P162 = PLOM_UNIFORM64(pv[162]);
    const double P163 = PLOM_UNIFORM64(pv[163]);
    const double P164 = PLOM_UNIFORM64(pv[164]);
    const double P165 = PLOM_UNIFORM64(pv[165]);
    const double P166 = PLOM_UNIFORM64(pv[166]);
    const double P167 = PLOM_UNIFORM64(pv[167]);
    const double P168 = PLOM_UNIFORM64(pv[168]);
    const double P169 = PLOM_UNIFORM64(pv[169]);
    const double P170 = PLOM_UNIFORM64(pv[170]);
    const double P171 = PLOM_UNIFORM64(pv[171]);
    const double P172 = PLOM_UNIFORM64(pv[172]);
    const double P173 = PLOM_UNIFORM64(pv[173]);
    const double P174 = PLOM_UNIFORM64(pv[174]);
    const double P175 = PLOM_UNIFORM64(pv[175]);
    const double P176 = PLOM_UNIFORM64(pv[176]);
    const double P177 = PLOM_UNIFORM64(pv[177]);
    const double P178 = PLOM_UNIFORM64(pv[178]);
    const double P179 = PLOM_UNIFORM64(pv[179]);
    X[30] = 0.0; /* reset_inc */
    X[31] = 0.0; /* reset_inc */
    X[32] = 0.0; /* reset_inc */
    X[33] = 0.0; /* reset_inc */
    X[34] = 0.0; /* reset_inc */
    X[35] = 0.0; /* reset_inc */
    X[36] = 0.0; /* reset_inc */
    X[37] = 0.0; /* reset_inc */
    X[38] = 0.0; /* reset_inc */
    X[39] = 0.0; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const double tr = (double)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const double C0 = plom_cov_ld(cvk + 0);
        const double C1 = plom_cov_ld(cvk + 1);
        const double C2 = plom_cov_ld(cvk + 2);
        const double C3 = plom_cov_ld(cvk + 3);
        const double C4 = plom_cov_ld(cvk + 4);
        const double C5 = plom_cov_ld(cvk + 5);
        const double C6 = plom_cov_ld(cvk + 6);
        const double C7 = plom_cov_ld(cvk + 7);
        const double C8 = plom_cov_ld(cvk + 8);
        const double C9 = plom_cov_ld(cvk + 9);
        double s30 = 0.0;
        double s31 = 0.0;
        double s32 = 0.0;
        double s33 = 0.0;
        double s34 = 0.0;
        double s35 = 0.0;
        double s36 = 0.0;
        double s37 = 0.0;
        double s38 = 0.0;
        double s39 = 0.0;
        const double r0_0 = ((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])) / C0);
        const double r1_0 = P10;
        const double r2_0 = P20;
        const double r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        const double f0 = X[0] + ( - (r0_0*X[0]) + (r3_0))*dtr;
        const double f10 = X[10] + ( - (r1_0*X[10]) + (r0_0*X[0]))*dtr;
        const double f20 = X[20] + ( - (r2_0*X[20]) + (r1_0*X[10]))*dtr;
        s30 += ( + (r1_0*X[10]))*dtr;
        X[0] = (f0 < 0.0) ? 0.0 : f0;
        X[10] = (f10 < 0.0) ? 0.0 : f10;
        const double r0_1 = ((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])) / C1);
        const double r1_1 = P11;
        const double r2_1 = P21;
        const double r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        const double f1 = X[1] + ( - (r0_1*X[1]) + (r3_1))*dtr;
        const double f11 = X[11] + ( - (r1_1*X[11]) + (r0_1*X[1]))*dtr;
        const double f21 = X[21] + ( - (r2_1*X[21]) + (r1_1*X[11]))*dtr;
        s31 += ( + (r1_1*X[11]))*dtr;
        X[1] = (f1 < 0.0) ? 0.0 : f1;
        X[11] = (f11 < 0.0) ? 0.0 : f11;
        const double r0_2 = ((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])) / C2);
        const double r1_2 = P12;
        const double r2_2 = P22;
        const double r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        const double f2 = X[2] + ( - (r0_2*X[2]) + (r3_2))*dtr;
        const double f12 = X[12] + ( - (r1_2*X[12]) + (r0_2*X[2]))*dtr;
        const double f22 = X[22] + ( - (r2_2*X[22]) + (r1_2*X[12]))*dtr;
        s32 += ( + (r1_2*X[12]))*dtr;
        X[2] = (f2 < 0.0) ? 0.0 : f2;
        X[12] = (f12 < 0.0) ? 0.0 : f12;
        const double r0_3 = ((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])) / C3);
        const double r1_3 = P13;
        const double r2_3 = P23;
        const double r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        const double f3 = X[3] + ( - (r0_3*X[3]) + (r3_3))*dtr;
        const double f13 = X[13] + ( - (r1_3*X[13]) + (r0_3*X[3]))*dtr;
        const double f23 = X[23] + ( - (r2_3*X[23]) + (r1_3*X[13]))*dtr;
        s33 += ( + (r1_3*X[13]))*dtr;
        X[3] = (f3 < 0.0) ? 0.0 : f3;
        X[13] = (f13 < 0.0) ? 0.0 : f13;
        const double r0_4 = ((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])) / C4);
        const double r1_4 = P14;
        const double r2_4 = P24;
        const double r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        const double f4 = X[4] + ( - (r0_4*X[4]) + (r3_4))*dtr;
        const double f14 = X[14] + ( - (r1_4*X[14]) + (r0_4*X[4]))*dtr;
        const double f24 = X[24] + ( - (r2_4*X[24]) + (r1_4*X[14]))*dtr;
        s34 += ( + (r1_4*X[14]))*dtr;
        X[4] = (f4 < 0.0) ? 0.0 : f4;
        X[14] = (f14 < 0.0) ? 0.0 : f14;
        const double r0_5 = ((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])) / C5);
        const double r1_5 = P15;
        const double r2_5 = P25;
        const double r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        const double f5 = X[5] + ( - (r0_5*X[5]) + (r3_5))*dtr;
        const double f15 = X[15] + ( - (r1_5*X[15]) + (r0_5*X[5]))*dtr;
        const double f25 = X[25] + ( - (r2_5*X[25]) + (r1_5*X[15]))*dtr;
        s35 += ( + (r1_5*X[15]))*dtr;
        X[5] = (f5 < 0.0) ? 0.0 : f5;
        X[15] = (f15 < 0.0) ? 0.0 : f15;
        const double r0_6 = ((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])) / C6);
        const double r1_6 = P16;
        const double r2_6 = P26;
        const double r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        const double f6 = X[6] + ( - (r0_6*X[6]) + (r3_6))*dtr;
        const double f16 = X[16] + ( - (r1_6*X[16]) + (r0_6*X[6]))*dtr;
        const double f26 = X[26] + ( - (r2_6*X[26]) + (r1_6*X[16]))*dtr;
        s36 += ( + (r1_6*X[16]))*dtr;
        X[6] = (f6 < 0.0) ? 0.0 : f6;
        X[16] = (f16 < 0.0) ? 0.0 : f16;
        const double r0_7 = ((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])) / C7);
        const double r1_7 = P17;
        const double r2_7 = P27;
        const double r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        const double f7 = X[7] + ( - (r0_7*X[7]) + (r3_7))*dtr;
        const double f17 = X[17] + ( - (r1_7*X[17]) + (r0_7*X[7]))*dtr;
        const double f27 = X[27] + ( - (r2_7*X[27]) + (r1_7*X[17]))*dtr;
        s37 += ( + (r1_7*X[17]))*dtr;
        X[7] = (f7 < 0.0) ? 0.0 : f7;
        X[17] = (f17 < 0.0) ? 0.0 : f17;
        const double r0_8 = ((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])) / C8);
        const double r1_8 = P18;
        const double r2_8 = P28;
        const double r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        const double f8 = X[8] + ( - (r0_8*X[8]) + (r3_8))*dtr;
        const double f18 = X[18] + ( - (r1_8*X[18]) + (r0_8*X[8]))*dtr;
        const double f28 = X[28] + ( - (r2_8*X[28]) + (r1_8*X[18]))*dtr;
        s38 += ( + (r1_8*X[18]))*dtr;
        X[8] = (f8 < 0.0) ? 0.0 : f8;
        X[18] = (f18 < 0.0) ? 0.0 : f18;
        const double r0_9 = ((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])) / C9);
        const double r1_9 = P19;
        const double r2_9 = P29;
        const double r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        const double f9 = X[9] + ( - (r0_9*X[9]) + (r3_9))*dtr;
        const double f19 = X[19] + ( - (r1_9*X[19]) + (r0_9*X[9]))*dtr;
        const double f29 = X[29] + ( - (r2_9*X[29]) + (r1_9*X[19]))*dtr;
        s39 += ( + (r1_9*X[19]))*dtr;
        X[9] = (f9 < 0.0) ? 0.0 : f9;
        X[19] = (f19 < 0.0) ? 0.0 : f19;
        X[20] = (f20 < 0.0) ? 0.0 : f20;
        X[21] = (f21 < 0.0) ? 0.0 : f21;
        X[22] = (f22 < 0.0) ? 0.0 : f22;
        X[23] = (f23 < 0.0) ? 0.0 : f23;
        X[24] = (f24 < 0.0) ? 0.0 : f24;
        X[25] = (f25 < 0.0) ? 0.0 : f25;
        X[26] = (f26 < 0.0) ? 0.0 : f26;
        X[27] = (f27 < 0.0) ? 0.0 : f27;
        X[28] = (f28 < 0.0) ? 0.0 : f28;
        X[29] = (f29 < 0.0) ? 0.0 : f29;
        const double f30 = X[30] + s30;
        X[30] = (f30 < 0.0) ? 0.0 : f30;
        const double f31 = X[31] + s31;
        X[31] = (f31 < 0.0) ? 0.0 : f31;
        const double f32 = X[32] + s32;
        X[32] = (f32 < 0.0) ? 0.0 : f32;
        const double f33 = X[33] + s33;
        X[33] = (f33 < 0.0) ? 0.0 : f33;
        const double f34 = X[34] + s34;
        X[34] = (f34 < 0.0) ? 0.0 : f34;
        const double f35 = X[35] + s35;
        X[35] = (f35 < 0.0) ? 0.0 : f35;
        const double f36 = X[36] + s36;
        X[36] = (f36 < 0.0) ? 0.0 : f36;
        const double f37 = X[37] + s37;
        X[37] = (f37 < 0.0) ? 0.0 : f37;
        const double f38 = X[38] + s38;
        X[38] = (f38 < 0.0) ? 0.0 : f38;
        const double f39 = X[39] + s39;
        X[39] = (f39 < 0.0) ? 0.0 : f39;
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_proj2obs_pp_f64(const double *X, const double *pvd, const double *lpd, const double *cvd, double *obs)
{
    (void)pvd; (void)lpd; (void)cvd; (void)X;
    obs[0] = (double)X[30];
    obs[1] = (double)X[31];
    obs[2] = (double)X[32];
    obs[3] = (double)X[33];
    obs[4] = (double)X[34];
    obs[5] = (double)X[35];
    obs[6] = (double)X[36];
    obs[7] = (double)X[37];
    obs[8] = (double)X[38];
    obs[9] = (double)X[39];
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_interval_full_pp_f32(float *X, const float *pv, const float *lp, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
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
    const float P20 = lp[20];
    const float P21 = lp[21];
    const float P22 = lp[22];
    const float P23 = lp[23];
    const float P24 = lp[24];
    const float P25 = lp[25];
    const float P26 = lp[26];
    const float P27 = lp[27];
    const float P28 = lp[28];
    const float P29 = lp[29];
    const float P30 = lp[30];
    const float P31 = lp[31];
    const float P32 = lp[32];
    const float P33 = lp[33];
    const float P34 = lp[34];
    const float P35 = lp[35];
    const float P36 = lp[36];
    const float P37 = lp[37];
    const float P38 = lp[38];
    const float P39 = lp[39];
    const float P40 = lp[40];
    const float P41 = lp[41];
    const float P42 = lp[42];
    const float P43 = lp[43];
    const float P44 = lp[44];
    const float P45 = lp[45];
    const float P46 = lp[46];
    const float P47 = lp[47];
    const float P48 = lp[48];
    const float P49 = lp[49];
    const float P80 = PLOM_UNIFORM32(pv[80]);
    const float P81 = PLOM_UNIFORM32(pv[81]);
    const float P82 = PLOM_UNIFORM32(pv[82]);
    const float P83 = PLOM_UNIFORM32(pv[83]);
    const float P84 = PLOM_UNIFORM32(pv[84]);
    const float P85 = PLOM_UNIFORM32(pv[85]);
    const float P86 = PLOM_UNIFORM32(pv[86]);
    const float P87 = PLOM_UNIFORM32(pv[87]);
    const float P88 = PLOM_UNIFORM32(pv[88]);
    const float P89 = PLOM_UNIFORM32(pv[89]);
    const float P90 = PLOM_UNIFORM32(pv[90]);
    const float P91 = PLOM_UNIFORM32(pv[91]);
    const float P92 = PLOM_UNIFORM32(pv[92]);
    const float P93 = PLOM_UNIFORM32(pv[93]);
    const float P94 = PLOM_UNIFORM32(pv[94]);
    const float P95 = PLOM_UNIFORM32(pv[95]);
    const float P96 = PLOM_UNIFORM32(pv[96]);
    const float P97 = PLOM_UNIFORM32(pv[97]);
    const float P98 = PLOM_UNIFORM32(pv[98]);
    const float P99 = PLOM_UNIFORM32(pv[99]);
    const float P100 = PLOM_UNIFORM32(pv[100]);
    const float P101 = PLOM_UNIFORM32(pv[101]);
    const float P102 = PLOM_UNIFORM32(pv[102]);
    const float P103 = PLOM_UNIFORM32(pv[103]);
    const float P104 = PLOM_UNIFORM32(pv[104]);
    const float P105 = PLOM_UNIFORM32(pv[105]);
    const float P106 = PLOM_UNIFORM32(pv[106]);
    const float P107 = PLOM_UNIFORM32(pv[107]);
    const float P108 = PLOM_UNIFORM32(pv[108]);
    const float P109 = PLOM_UNIFORM32(pv[109]);
    const float P110 = PLOM_UNIFORM32(pv[110]);
    const float P111 = PLOM_UNIFORM32(pv[111]);
    const float P112 = PLOM_UNIFORM32(pv[112]);
    const float P113 = PLOM_UNIFORM32(pv[113]);
    const float P114 = PLOM_UNIFORM32(pv[114]);
    const float P115 = PLOM_UNIFORM32(pv[115]);
    const float P116 = PLOM_UNIFORM32(pv[116]);
    const float P117 = PLOM_UNIFORM32(pv[117]);
    const float P118 = PLOM_UNIFORM32(pv[118]);
    const float P119 = PLOM_UNIFORM32(pv[119]);
    const float P120 = PLOM_UNIFORM32(pv[120]);
    const float P121 = PLOM_UNIFORM32(pv[121]);
    const float P122 = PLOM_UNIFORM32(pv[122]);
    const float P123 = PLOM_UNIFORM32(pv[123]);
    const float P124 = PLOM_UNIFORM32(pv[124]);
    const float P125 = PLOM_UNIFORM32(pv[125]);
    const float P126 = PLOM_UNIFORM32(pv[126]);
    const float P127 = PLOM_UNIFORM32(pv[127]);
    const float P128 = PLOM_UNIFORM32(pv[128]);
    const float P129 = PLOM_UNIFORM32(pv[129]);
    const float P130 = PLOM_UNIFORM32(pv[130]);
    const float P131 = PLOM_UNIFORM32(pv[131]);
    const float P132 = PLOM_UNIFORM32(pv[132]);
    const float P133 = PLOM_UNIFORM32(pv[133]);
    const float P134 = PLOM_UNIFORM32(pv[134]);
    const float P135 = PLOM_UNIFORM32(pv[135]);
    const float P136 = PLOM_UNIFORM32(pv[136]);
    const float P137 = PLOM_UNIFORM32(pv[137]);
    const float P138 = PLOM_UNIFORM32(pv[138]);
    const float P139 = PLOM_UNIFORM32(pv[139]);
    const float P140 = PLOM_UNIFORM32(pv[140]);
    const float P141 = PLOM_UNIFORM32(pv[141]);
    const float P142 = PLOM_UNIFORM32(pv[142]);
    const float P143 = PLOM_UNIFORM32(pv[143]);
    const float P144 = PLOM_UNIFORM32(pv[144]);
    const float P145 = PLOM_UNIFORM32(pv[145]);
    const float P146 = PLOM_UNIFORM32(pv[146]);
    const float P147 = PLOM_UNIFORM32(pv[147]);
    const float P148 = PLOM_UNIFORM32(pv[148]);
    const float P149 = PLOM_UNIFORM32(pv[149]);
    const float P150 = PLOM_UNIFORM32(pv[150]);
    const float P151 = PLOM_UNIFORM32(pv[151]);
    const float P152 = PLOM_UNIFORM32(pv[152]);
    const float P153 = PLOM_UNIFORM32(pv[153]);
    const float P154 = PLOM_UNIFORM32(pv[154]);
    const float P155 = PLOM_UNIFORM32(pv[155]);
    const float P156 = PLOM_UNIFORM32(pv[156]);
    const float P157 = PLOM_UNIFORM32(pv[157]);
    const float P158 = PLOM_UNIFORM32(pv[158]);
    const float P159 = PLOM_UNIFORM32(pv[159]);
    const float P160 = PLOM_UNIFORM32(pv[160]);
    const float P161 = PLOM_UNIFORM32(pv[161]);
    const float P162 = PLOM_UNIFORM32(pv[162]);
    const float P163 = PLOM_UNIFORM32(pv[163]);
    const float P164 = PLOM_UNIFORM32(pv[164]);
    const float P165 = PLOM_UNIFORM32(pv[165]);
    const float P166 = PLOM_UNIFORM32(pv[166]);
    const float P167 = PLOM_UNIFORM32(pv[167]);
    const float P168 = PLOM_UNIFORM32(pv[168]);
    const float P169 = PLOM_UNIFORM32(pv[169]);
    const float P170 = PLOM_UNIFORM32(pv[170]);
    const float P171 = PLOM_UNIFORM32(pv[171]);
    const float P172 = PLOM_UNIFORM32(pv[172]);
    const float P173 = PLOM_UNIFORM32(pv[173]);
    const float P174 = PLOM_UNIFORM32(pv[174]);
    const float P175 = PLOM_UNIFORM32(pv[175]);
    const float P176 = PLOM_UNIFORM32(pv[176]);
    const float P177 = PLOM_UNIFORM32(pv[177]);
    const float P178 = PLOM_UNIFORM32(pv[178]);
    const float P179 = PLOM_UNIFORM32(pv[179]);
    X[30] = 0.0f; /* reset_inc */
    X[31] = 0.0f; /* reset_inc */
    X[32] = 0.0f; /* reset_inc */
    X[33] = 0.0f; /* reset_inc */
    X[34] = 0.0f; /* reset_inc */
    X[35] = 0.0f; /* reset_inc */
    X[36] = 0.0f; /* reset_inc */
    X[37] = 0.0f; /* reset_inc */
    X[38] = 0.0f; /* reset_inc */
    X[39] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        const float C1 = (float)plom_cov_ld(cvk + 1);
        const float C2 = (float)plom_cov_ld(cvk + 2);
        const float C3 = (float)plom_cov_ld(cvk + 3);
        const float C4 = (float)plom_cov_ld(cvk + 4);
        const float C5 = (float)plom_cov_ld(cvk + 5);
        const float C6 = (float)plom_cov_ld(cvk + 6);
        const float C7 = (float)plom_cov_ld(cvk + 7);
        const float C8 = (float)plom_cov_ld(cvk + 8);
        const float C9 = (float)plom_cov_ld(cvk + 9);
        float s30 = 0.0f;
        float s31 = 0.0f;
        float s32 = 0.0f;
        float s33 = 0.0f;
        float s34 = 0.0f;
        float s35 = 0.0f;
        float s36 = 0.0f;
        float s37 = 0.0f;
        float s38 = 0.0f;
        float s39 = 0.0f;
        const float r0_0 = plom_fdivf((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])), C0);
        const float r1_0 = P10;
        const float r2_0 = P20;
        const float r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        float za0, zb0;
        plom_normal_pair32(ns, &za0, &zb0);
        const float d0_0 = sqdt * za0;
        const float d1_0 = sqdt * zb0;
        float za1, zb1;
        plom_normal_pair32(ns, &za1, &zb1);
        const float d2_0 = sqdt * za1;
        const float d3_0 = sqdt * zb1;
        float za2, zb2;
        plom_normal_pair32(ns, &za2, &zb2);
        const float w0_0 = sqdt * za2;
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r3_0))*dtr + - plom_fsqrtf((r0_0*X[0]))*d0_0+ plom_fsqrtf((r3_0))*d3_0 + - (r0_0*X[0])*P40*w0_0;
        const float f10 = X[10] + ( - (r1_0*X[10]) + (r0_0*X[0]))*dtr + - plom_fsqrtf((r1_0*X[10]))*d1_0+ plom_fsqrtf((r0_0*X[0]))*d0_0 + + (r0_0*X[0])*P40*w0_0;
        const float f20 = X[20] + ( - (r2_0*X[20]) + (r1_0*X[10]))*dtr + - plom_fsqrtf((r2_0*X[20]))*d2_0+ plom_fsqrtf((r1_0*X[10]))*d1_0;
        s30 += ( + (r1_0*X[10]))*dtr + + plom_fsqrtf((r1_0*X[10]))*d1_0;
        X[0] = plom_clipf(f0);
        X[10] = plom_clipf(f10);
        const float r0_1 = plom_fdivf((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])), C1);
        const float r1_1 = P11;
        const float r2_1 = P21;
        const float r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        const float d0_1 = sqdt * zb2;
        float za3, zb3;
        plom_normal_pair32(ns, &za3, &zb3);
        const float d1_1 = sqdt * za3;
        const float d2_1 = sqdt * zb3;
        float za4, zb4;
        plom_normal_pair32(ns, &za4, &zb4);
        const float d3_1 = sqdt * za4;
        const float w0_1 = sqdt * zb4;
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r3_1))*dtr + - plom_fsqrtf((r0_1*X[1]))*d0_1+ plom_fsqrtf((r3_1))*d3_1 + - (r0_1*X[1])*P41*w0_1;
        const float f11 = X[11] + ( - (r1_1*X[11]) + (r0_1*X[1]))*dtr + - plom_fsqrtf((r1_1*X[11]))*d1_1+ plom_fsqrtf((r0_1*X[1]))*d0_1 + + (r0_1*X[1])*P41*w0_1;
        const float f21 = X[21] + ( - (r2_1*X[21]) + (r1_1*X[11]))*dtr + - plom_fsqrtf((r2_1*X[21]))*d2_1+ plom_fsqrtf((r1_1*X[11]))*d1_1;
        s31 += ( + (r1_1*X[11]))*dtr + + plom_fsqrtf((r1_1*X[11]))*d1_1;
        X[1] = plom_clipf(f1);
        X[11] = plom_clipf(f11);
        const float r0_2 = plom_fdivf((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])), C2);
        const float r1_2 = P12;
        const float r2_2 = P22;
        const float r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        float za5, zb5;
        plom_normal_pair32(ns, &za5, &zb5);
        const float d0_2 = sqdt * za5;
        const float d1_2 = sqdt * zb5;
        float za6, zb6;
        plom_normal_pair32(ns, &za6, &zb6);
        const float d2_2 = sqdt * za6;
        const float d3_2 = sqdt * zb6;
        float za7, zb7;
        plom_normal_pair32(ns, &za7, &zb7);
        const float w0_2 = sqdt * za7;
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r3_2))*dtr + - plom_fsqrtf((r0_2*X[2]))*d0_2+ plom_fsqrtf((r3_2))*d3_2 + - (r0_2*X[2])*P42*w0_2;
        const float f12 = X[12] + ( - (r1_2*X[12]) + (r0_2*X[2]))*dtr + - plom_fsqrtf((r1_2*X[12]))*d1_2+ plom_fsqrtf((r0_2*X[2]))*d0_2 + + (r0_2*X[2])*P42*w0_2;
        const float f22 = X[22] + ( - (r2_2*X[22]) + (r1_2*X[12]))*dtr + - plom_fsqrtf((r2_2*X[22]))*d2_2+ plom_fsqrtf((r1_2*X[12]))*d1_2;
        s32 += ( + (r1_2*X[12]))*dtr + + plom_fsqrtf((r1_2*X[12]))*d1_2;
        X[2] = plom_clipf(f2);
        X[12] = plom_clipf(f12);
        const float r0_3 = plom_fdivf((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])), C3);
        const float r1_3 = P13;
        const float r2_3 = P23;
        const float r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        const float d0_3 = sqdt * zb7;
        float za8, zb8;
        plom_normal_pair32(ns, &za8, &zb8);
        const float d1_3 = sqdt * za8;
        const float d2_3 = sqdt * zb8;
        float za9, zb9;
        plom_normal_pair32(ns, &za9, &zb9);
        const float d3_3 = sqdt * za9;
        const float w0_3 = sqdt * zb9;
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r3_3))*dtr + - plom_fsqrtf((r0_3*X[3]))*d0_3+ plom_fsqrtf((r3_3))*d3_3 + - (r0_3*X[3])*P43*w0_3;
        const float f13 = X[13] + ( - (r1_3*X[13]) + (r0_3*X[3]))*dtr + - plom_fsqrtf((r1_3*X[13]))*d1_3+ plom_fsqrtf((r0_3*X[3]))*d0_3 + + (r0_3*X[3])*P43*w0_3;
        const float f23 = X[23] + ( - (r2_3*X[23]) + (r1_3*X[13]))*dtr + - plom_fsqrtf((r2_3*X[23]))*d2_3+ plom_fsqrtf((r1_3*X[13]))*d1_3;
        s33 += ( + (r1_3*X[13]))*dtr + + plom_fsqrtf((r1_3*X[13]))*d1_3;
        X[3] = plom_clipf(f3);
        X[13] = plom_clipf(f13);
        const float r0_4 = plom_fdivf((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])), C4);
        const float r1_4 = P14;
        const float r2_4 = P24;
        const float r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        float za10, zb10;
        plom_normal_pair32(ns, &za10, &zb10);
        const float d0_4 = sqdt * za10;
        const float d1_4 = sqdt * zb10;
        float za11, zb11;
        plom_normal_pair32(ns, &za11, &zb11);
        const float d2_4 = sqdt * za11;
        const float d3_4 = sqdt * zb11;
        float za12, zb12;
        plom_normal_pair32(ns, &za12, &zb12);
        const float w0_4 = sqdt * za12;
        const float f4 = X[4] + ( - (r0_4*X[4]) + (r3_4))*dtr + - plom_fsqrtf((r0_4*X[4]))*d0_4+ plom_fsqrtf((r3_4))*d3_4 + - (r0_4*X[4])*P44*w0_4;
        const float f14 = X[14] + ( - (r1_4*X[14]) + (r0_4*X[4]))*dtr + - plom_fsqrtf((r1_4*X[14]))*d1_4+ plom_fsqrtf((r0_4*X[4]))*d0_4 + + (r0_4*X[4])*P44*w0_4;
        const float f24 = X[24] + ( - (r2_4*X[24]) + (r1_4*X[14]))*dtr + - plom_fsqrtf((r2_4*X[24]))*d2_4+ plom_fsqrtf((r1_4*X[14]))*d1_4;
        s34 += ( + (r1_4*X[14]))*dtr + + plom_fsqrtf((r1_4*X[14]))*d1_4;
        X[4] = plom_clipf(f4);
        X[14] = plom_clipf(f14);
        const float r0_5 = plom_fdivf((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])), C5);
        const float r1_5 = P15;
        const float r2_5 = P25;
        const float r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        const float d0_5 = sqdt * zb12;
        float za13, zb13;
        plom_normal_pair32(ns, &za13, &zb13);
        const float d1_5 = sqdt * za13;
        const float d2_5 = sqdt * zb13;
        float za14, zb14;
        plom_normal_pair32(ns, &za14, &zb14);
        const float d3_5 = sqdt * za14;
        const float w0_5 = sqdt * zb14;
        const float f5 = X[5] + ( - (r0_5*X[5]) + (r3_5))*dtr + - plom_fsqrtf((r0_5*X[5]))*d0_5+ plom_fsqrtf((r3_5))*d3_5 + - (r0_5*X[5])*P45*w0_5;
        const float f15 = X[15] + ( - (r1_5*X[15]) + (r0_5*X[5]))*dtr + - plom_fsqrtf((r1_5*X[15]))*d1_5+ plom_fsqrtf((r0_5*X[5]))*d0_5 + + (r0_5*X[5])*P45*w0_5;
        const float f25 = X[25] + ( - (r2_5*X[25]) + (r1_5*X[15]))*dtr + - plom_fsqrtf((r2_5*X[25]))*d2_5+ plom_fsqrtf((r1_5*X[15]))*d1_5;
        s35 += ( + (r1_5*X[15]))*dtr + + plom_fsqrtf((r1_5*X[15]))*d1_5;
        X[5] = plom_clipf(f5);
        X[15] = plom_clipf(f15);
        const float r0_6 = plom_fdivf((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])), C6);
        const float r1_6 = P16;
        const float r2_6 = P26;
        const float r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        float za15, zb15;
        plom_normal_pair32(ns, &za15, &zb15);
        const float d0_6 = sqdt * za15;
        const float d1_6 = sqdt * zb15;
        float za16, zb16;
        plom_normal_pair32(ns, &za16, &zb16);
        const float d2_6 = sqdt * za16;
        const float d3_6 = sqdt * zb16;
        float za17, zb17;
        plom_normal_pair32(ns, &za17, &zb17);
        const float w0_6 = sqdt * za17;
        const float f6 = X[6] + ( - (r0_6*X[6]) + (r3_6))*dtr + - plom_fsqrtf((r0_6*X[6]))*d0_6+ plom_fsqrtf((r3_6))*d3_6 + - (r0_6*X[6])*P46*w0_6;
        const float f16 = X[16] + ( - (r1_6*X[16]) + (r0_6*X[6]))*dtr + - plom_fsqrtf((r1_6*X[16]))*d1_6+ plom_fsqrtf((r0_6*X[6]))*d0_6 + + (r0_6*X[6])*P46*w0_6;
        const float f26 = X[26] + ( - (r2_6*X[26]) + (r1_6*X[16]))*dtr + - plom_fsqrtf((r2_6*X[26]))*d2_6+ plom_fsqrtf((r1_6*X[16]))*d1_6;
        s36 += ( + (r1_6*X[16]))*dtr + + plom_fsqrtf((r1_6*X[16]))*d1_6;
        X[6] = plom_clipf(f6);
        X[16] = plom_clipf(f16);
        const float r0_7 = plom_fdivf((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])), C7);
        const float r1_7 = P17;
        const float r2_7 = P27;
        const float r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        const float d0_7 = sqdt * zb17;
        float za18, zb18;
        plom_normal_pair32(ns, &za18, &zb18);
        const float d1_7 = sqdt * za18;
        const float d2_7 = sqdt * zb18;
        float za19, zb19;
        plom_normal_pair32(ns, &za19, &zb19);
        const float d3_7 = sqdt * za19;
        const float w0_7 = sqdt * zb19;
        const float f7 = X[7] + ( - (r0_7*X[7]) + (r3_7))*dtr + - plom_fsqrtf((r0_7*X[7]))*d0_7+ plom_fsqrtf((r3_7))*d3_7 + - (r0_7*X[7])*P47*w0_7;
        const float f17 = X[17] + ( - (r1_7*X[17]) + (r0_7*X[7]))*dtr + - plom_fsqrtf((r1_7*X[17]))*d1_7+ plom_fsqrtf((r0_7*X[7]))*d0_7 + + (r0_7*X[7])*P47*w0_7;
        const float f27 = X[27] + ( - (r2_7*X[27]) + (r1_7*X[17]))*dtr + - plom_fsqrtf((r2_7*X[27]))*d2_7+ plom_fsqrtf((r1_7*X[17]))*d1_7;
        s37 += ( + (r1_7*X[17]))*dtr + + plom_fsqrtf((r1_7*X[17]))*d1_7;
        X[7] = plom_clipf(f7);
        X[17] = plom_clipf(f17);
        const float r0_8 = plom_fdivf((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])), C8);
        const float r1_8 = P18;
        const float r2_8 = P28;
        const float r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        float za20, zb20;
        plom_normal_pair32(ns, &za20, &zb20);
        const float d0_8 = sqdt * za20;
        const float d1_8 = sqdt * zb20;
        float za21, zb21;
        plom_normal_pair32(ns, &za21, &zb21);
        const float d2_8 = sqdt * za21;
        const float d3_8 = sqdt * zb21;
        float za22, zb22;
        plom_normal_pair32(ns, &za22, &zb22);
        const float w0_8 = sqdt * za22;
        const float f8 = X[8] + ( - (r0_8*X[8]) + (r3_8))*dtr + - plom_fsqrtf((r0_8*X[8]))*d0_8+ plom_fsqrtf((r3_8))*d3_8 + - (r0_8*X[8])*P48*w0_8;
        const float f18 = X[18] + ( - (r1_8*X[18]) + (r0_8*X[8]))*dtr + - plom_fsqrtf((r1_8*X[18]))*d1_8+ plom_fsqrtf((r0_8*X[8]))*d0_8 + + (r0_8*X[8])*P48*w0_8;
        const float f28 = X[28] + ( - (r2_8*X[28]) + (r1_8*X[18]))*dtr + - plom_fsqrtf((r2_8*X[28]))*d2_8+ plom_fsqrtf((r1_8*X[18]))*d1_8;
        s38 += ( + (r1_8*X[18]))*dtr + + plom_fsqrtf((r1_8*X[18]))*d1_8;
        X[8] = plom_clipf(f8);
        X[18] = plom_clipf(f18);
        const float r0_9 = plom_fdivf((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])), C9);
        const float r1_9 = P19;
        const float r2_9 = P29;
        const float r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        const float d0_9 = sqdt * zb22;
        float za23, zb23;
        plom_normal_pair32(ns, &za23, &zb23);
        const float d1_9 = sqdt * za23;
        const float d2_9 = sqdt * zb23;
        float za24, zb24;
        plom_normal_pair32(ns, &za24, &zb24);
        const float d3_9 = sqdt * za24;
        const float w0_9 = sqdt * zb24;
        const float f9 = X[9] + ( - (r0_9*X[9]) + (r3_9))*dtr + - plom_fsqrtf((r0_9*X[9]))*d0_9+ plom_fsqrtf((r3_9))*d3_9 + - (r0_9*X[9])*P49*w0_9;
        const float f19 = X[19] + ( - (r1_9*X[19]) + (r0_9*X[9]))*dtr + - plom_fsqrtf((r1_9*X[19]))*d1_9+ plom_fsqrtf((r0_9*X[9]))*d0_9 + + (r0_9*X[9])*P49*w0_9;
        const float f29 = X[29] + ( - (r2_9*X[29]) + (r1_9*X[19]))*dtr + - plom_fsqrtf((r2_9*X[29]))*d2_9+ plom_fsqrtf((r1_9*X[19]))*d1_9;
        s39 += ( + (r1_9*X[19]))*dtr + + plom_fsqrtf((r1_9*X[19]))*d1_9;
        X[9] = plom_clipf(f9);
        X[19] = plom_clipf(f19);
        X[20] = plom_clipf(f20);
        X[21] = plom_clipf(f21);
        X[22] = plom_clipf(f22);
        X[23] = plom_clipf(f23);
        X[24] = plom_clipf(f24);
        X[25] = plom_clipf(f25);
        X[26] = plom_clipf(f26);
        X[27] = plom_clipf(f27);
        X[28] = plom_clipf(f28);
        X[29] = plom_clipf(f29);
        const float f30 = X[30] + s30;
        X[30] = plom_clipf(f30);
        const float f31 = X[31] + s31;
        X[31] = plom_clipf(f31);
        const float f32 = X[32] + s32;
        X[32] = plom_clipf(f32);
        const float f33 = X[33] + s33;
        X[33] = plom_clipf(f33);
        const float f34 = X[34] + s34;
        X[34] = plom_clipf(f34);
        const float f35 = X[35] + s35;
        X[35] = plom_clipf(f35);
        const float f36 = X[36] + s36;
        X[36] = plom_clipf(f36);
        const float f37 = X[37] + s37;
        X[37] = plom_clipf(f37);
        const float f38 = X[38] + s38;
        X[38] = plom_clipf(f38);
        const float f39 = X[39] + s39;
        X[39] = plom_clipf(f39);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_interval_no_dem_sto_pp_f32(float *X, const float *pv, const float *lp, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
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
    const float P20 = lp[20];
    const float P21 = lp[21];
    const float P22 = lp[22];
    const float P23 = lp[23];
    const float P24 = lp[24];
    const float P25 = lp[25];
    const float P26 = lp[26];
    const float P27 = lp[27];
    const float P28 = lp[28];
    const float P29 = lp[29];
    const float P30 = lp[30];
    const float P31 = lp[31];
    const float P32 = lp[32];
    const float P33 = lp[33];
    const float P34 = lp[34];
    const float P35 = lp[35];
    const float P36 = lp[36];
    const float P37 = lp[37];
    const float P38 = lp[38];
    const float P39 = lp[39];
    const float P40 = lp[40];
    const float P41 = lp[41];
    const float P42 = lp[42];
    const float P43 = lp[43];
    const float P44 = lp[44];
    const float P45 = lp[45];
    const float P46 = lp[46];
    const float P47 = lp[47];
    const float P48 = lp[48];
    const float P49 = lp[49];
    const float P80 = PLOM_UNIFORM32(pv[80]);
    const float P81 = PLOM_UNIFORM32(pv[81]);
    const float P82 = PLOM_UNIFORM32(pv[82]);
    const float P83 = PLOM_UNIFORM32(pv[83]);
    const float P84 = PLOM_UNIFORM32(pv[84]);
    const float P85 = PLOM_UNIFORM32(pv[85]);
    const float P86 = PLOM_UNIFORM32(pv[86]);
    const float P87 = PLOM_UNIFORM32(pv[87]);
    const float P88 = PLOM_UNIFORM32(pv[88]);
    const float P89 = PLOM_UNIFORM32(pv[89]);
    const float P90 = PLOM_UNIFORM32(pv[90]);
    const float P91 = PLOM_UNIFORM32(pv[91]);
    const float P92 = PLOM_UNIFORM32(pv[92]);
    const float P93 = PLOM_UNIFORM32(pv[93]);
    const float P94 = PLOM_UNIFORM32(pv[94]);
    const float P95 = PLOM_UNIFORM32(pv[95]);
    const float P96 = PLOM_UNIFORM32(pv[96]);
    const float P97 = PLOM_UNIFORM32(pv[97]);
    const float P98 = PLOM_UNIFORM32(pv[98]);
    const float P99 = PLOM_UNIFORM32(pv[99]);
    const float P100 = PLOM_UNIFORM32(pv[100]);
    const float P101 = PLOM_UNIFORM32(pv[101]);
    const float P102 = PLOM_UNIFORM32(pv[102]);
    const float P103 = PLOM_UNIFORM32(pv[103]);
    const float P104 = PLOM_UNIFORM32(pv[104]);
    const float P105 = PLOM_UNIFORM32(pv[105]);
    const float P106 = PLOM_UNIFORM32(pv[106]);
    const float P107 = PLOM_UNIFORM32(pv[107]);
    const float P108 = PLOM_UNIFORM32(pv[108]);
    const float P109 = PLOM_UNIFORM32(pv[109]);
    const float P110 = PLOM_UNIFORM32(pv[110]);
    const float P111 = PLOM_UNIFORM32(pv[111]);
    const float P112 = PLOM_UNIFORM32(pv[112]);
    const float P113 = PLOM_UNIFORM32(pv[113]);
    const float P114 = PLOM_UNIFORM32(pv[114]);
    const float P115 = PLOM_UNIFORM32(pv[115]);
    const float P116 = PLOM_UNIFORM32(pv[116]);
    const float P117 = PLOM_UNIFORM32(pv[117]);
    const float P118 = PLOM_UNIFORM32(pv[118]);
    const float P119 = PLOM_UNIFORM32(pv[119]);
    const float P120 = PLOM_UNIFORM32(pv[120]);
    const float P121 = PLOM_UNIFORM32(pv[121]);
    const float P122 = PLOM_UNIFORM32(pv[122]);
    const float P123 = PLOM_UNIFORM32(pv[123]);
    const float P124 = PLOM_UNIFORM32(pv[124]);
    const float P125 = PLOM_UNIFORM32(pv[125]);
    const float P126 = PLOM_UNIFORM32(pv[126]);
    const float P127 = PLOM_UNIFORM32(pv[127]);
    const float P128 = PLOM_UNIFORM32(pv[128]);
    const float P129 = PLOM_UNIFORM32(pv[129]);
    const float P130 = PLOM_UNIFORM32(pv[130]);
    const float P131 = PLOM_UNIFORM32(pv[131]);
    const float P132 = PLOM_UNIFORM32(pv[132]);
    const float P133 = PLOM_UNIFORM32(pv[133]);
    const float P134 = PLOM_UNIFORM32(pv[134]);
    const float P135 = PLOM_UNIFORM32(pv[135]);
    const float P136 = PLOM_UNIFORM32(pv[136]);
    const float P137 = PLOM_UNIFORM32(pv[137]);
    const float P138 = PLOM_UNIFORM32(pv[138]);
    const float P139 = PLOM_UNIFORM32(pv[139]);
    const float P140 = PLOM_UNIFORM32(pv[140]);
    const float P141 = PLOM_UNIFORM32(pv[141]);
    const float P142 = PLOM_UNIFORM32(pv[142]);
    const float P143 = PLOM_UNIFORM32(pv[143]);
    const float P144 = PLOM_UNIFORM32(pv[144]);
    const float P145 = PLOM_UNIFORM32(pv[145]);
    const float P146 = PLOM_UNIFORM32(pv[146]);
    const float P147 = PLOM_UNIFORM32(pv[147]);
    const float P148 = PLOM_UNIFORM32(pv[148]);
    const float P149 = PLOM_UNIFORM32(pv[149]);
    const float P150 = PLOM_UNIFORM32(pv[150]);
    const float P151 = PLOM_UNIFORM32(pv[151]);
    const float P152 = PLOM_UNIFORM32(pv[152]);
    const float P153 = PLOM_UNIFORM32(pv[153]);
    const float P154 = PLOM_UNIFORM32(pv[154]);
    const float P155 = PLOM_UNIFORM32(pv[155]);
    const float P156 = PLOM_UNIFORM32(pv[156]);
    const float P157 = PLOM_UNIFORM32(pv[157]);
    const float P158 = PLOM_UNIFORM32(pv[158]);
    const float P159 = PLOM_UNIFORM32(pv[159]);
    const float P160 = PLOM_UNIFORM32(pv[160]);
    const float P161 = PLOM_UNIFORM32(pv[161]);
    const float P162 = PLOM_UNIFORM32(pv[162]);
    const float P163 = PLOM_UNIFORM32(pv[163]);
    const float P164 = PLOM_UNIFORM32(pv[164]);
    const float P165 = PLOM_UNIFORM32(pv[165]);
    const float P166 = PLOM_UNIFORM32(pv[166]);
    const float P167 = PLOM_UNIFORM32(pv[167]);
    const float P168 = PLOM_UNIFORM32(pv[168]);
    const float P169 = PLOM_UNIFORM32(pv[169]);
    const float P170 = PLOM_UNIFORM32(pv[170]);
    const float P171 = PLOM_UNIFORM32(pv[171]);
    const float P172 = PLOM_UNIFORM32(pv[172]);
    const float P173 = PLOM_UNIFORM32(pv[173]);
    const float P174 = PLOM_UNIFORM32(pv[174]);
    const float P175 = PLOM_UNIFORM32(pv[175]);
    const float P176 = PLOM_UNIFORM32(pv[176]);
    const float P177 = PLOM_UNIFORM32(pv[177]);
    const float P178 = PLOM_UNIFORM32(pv[178]);
    const float P179 = PLOM_UNIFORM32(pv[179]);
    X[30] = 0.0f; /* reset_inc */
    X[31] = 0.0f; /* reset_inc */
    X[32] = 0.0f; /* reset_inc */
    X[33] = 0.0f; /* reset_inc */
    X[34] = 0.0f; /* reset_inc */
    X[35] = 0.0f; /* reset_inc */
    X[36] = 0.0f; /* reset_inc */
    X[37] = 0.0f; /* reset_inc */
    X[38] = 0.0f; /* reset_inc */
    X[39] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        const float C1 = (float)plom_cov_ld(cvk + 1);
        const float C2 = (float)plom_cov_ld(cvk + 2);
        const float C3 = (float)plom_cov_ld(cvk + 3);
        const float C4 = (float)plom_cov_ld(cvk + 4);
        const float C5 = (float)plom_cov_ld(cvk + 5);
        const float C6 = (float)plom_cov_ld(cvk + 6);
        const float C7 = (float)plom_cov_ld(cvk + 7);
        const float C8 = (float)plom_cov_ld(cvk + 8);
        const float C9 = (float)plom_cov_ld(cvk + 9);
        float s30 = 0.0f;
        float s31 = 0.0f;
        float s32 = 0.0f;
        float s33 = 0.0f;
        float s34 = 0.0f;
        float s35 = 0.0f;
        float s36 = 0.0f;
        float s37 = 0.0f;
        float s38 = 0.0f;
        float s39 = 0.0f;
        const float r0_0 = plom_fdivf((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])), C0);
        const float r1_0 = P10;
        const float r2_0 = P20;
        const float r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        float za0, zb0;
        plom_normal_pair32(ns, &za0, &zb0);
        const float w0_0 = sqdt * za0;
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r3_0))*dtr + - (r0_0*X[0])*P40*w0_0;
        const float f10 = X[10] + ( - (r1_0*X[10]) + (r0_0*X[0]))*dtr + + (r0_0*X[0])*P40*w0_0;
        const float f20 = X[20] + ( - (r2_0*X[20]) + (r1_0*X[10]))*dtr;
        s30 += ( + (r1_0*X[10]))*dtr;
        X[0] = plom_clipf(f0);
        X[10] = plom_clipf(f10);
        const float r0_1 = plom_fdivf((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])), C1);
        const float r1_1 = P11;
        const float r2_1 = P21;
        const float r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        const float w0_1 = sqdt * zb0;
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r3_1))*dtr + - (r0_1*X[1])*P41*w0_1;
        const float f11 = X[11] + ( - (r1_1*X[11]) + (r0_1*X[1]))*dtr + + (r0_1*X[1])*P41*w0_1;
        const float f21 = X[21] + ( - (r2_1*X[21]) + (r1_1*X[11]))*dtr;
        s31 += ( + (r1_1*X[11]))*dtr;
        X[1] = plom_clipf(f1);
        X[11] = plom_clipf(f11);
        const float r0_2 = plom_fdivf((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])), C2);
        const float r1_2 = P12;
        const float r2_2 = P22;
        const float r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        float za1, zb1;
        plom_normal_pair32(ns, &za1, &zb1);
        const float w0_2 = sqdt * za1;
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r3_2))*dtr + - (r0_2*X[2])*P42*w0_2;
        const float f12 = X[12] + ( - (r1_2*X[12]) + (r0_2*X[2]))*dtr + + (r0_2*X[2])*P42*w0_2;
        const float f22 = X[22] + ( - (r2_2*X[22]) + (r1_2*X[12]))*dtr;
        s32 += ( + (r1_2*X[12]))*dtr;
        X[2] = plom_clipf(f2);
        X[12] = plom_clipf(f12);
        const float r0_3 = plom_fdivf((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])), C3);
        const float r1_3 = P13;
        const float r2_3 = P23;
        const float r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        const float w0_3 = sqdt * zb1;
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r3_3))*dtr + - (r0_3*X[3])*P43*w0_3;
        const float f13 = X[13] + ( - (r1_3*X[13]) + (r0_3*X[3]))*dtr + + (r0_3*X[3])*P43*w0_3;
        const float f23 = X[23] + ( - (r2_3*X[23]) + (r1_3*X[13]))*dtr;
        s33 += ( + (r1_3*X[13]))*dtr;
        X[3] = plom_clipf(f3);
        X[13] = plom_clipf(f13);
        const float r0_4 = plom_fdivf((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])), C4);
        const float r1_4 = P14;
        const float r2_4 = P24;
        const float r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        float za2, zb2;
        plom_normal_pair32(ns, &za2, &zb2);
        const float w0_4 = sqdt * za2;
        const float f4 = X[4] + ( - (r0_4*X[4]) + (r3_4))*dtr + - (r0_4*X[4])*P44*w0_4;
        const float f14 = X[14] + ( - (r1_4*X[14]) + (r0_4*X[4]))*dtr + + (r0_4*X[4])*P44*w0_4;
        const float f24 = X[24] + ( - (r2_4*X[24]) + (r1_4*X[14]))*dtr;
        s34 += ( + (r1_4*X[14]))*dtr;
        X[4] = plom_clipf(f4);
        X[14] = plom_clipf(f14);
        const float r0_5 = plom_fdivf((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])), C5);
        const float r1_5 = P15;
        const float r2_5 = P25;
        const float r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        const float w0_5 = sqdt * zb2;
        const float f5 = X[5] + ( - (r0_5*X[5]) + (r3_5))*dtr + - (r0_5*X[5])*P45*w0_5;
        const float f15 = X[15] + ( - (r1_5*X[15]) + (r0_5*X[5]))*dtr + + (r0_5*X[5])*P45*w0_5;
        const float f25 = X[25] + ( - (r2_5*X[25]) + (r1_5*X[15]))*dtr;
        s35 += ( + (r1_5*X[15]))*dtr;
        X[5] = plom_clipf(f5);
        X[15] = plom_clipf(f15);
        const float r0_6 = plom_fdivf((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])), C6);
        const float r1_6 = P16;
        const float r2_6 = P26;
        const float r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        float za3, zb3;
        plom_normal_pair32(ns, &za3, &zb3);
        const float w0_6 = sqdt * za3;
        const float f6 = X[6] + ( - (r0_6*X[6]) + (r3_6))*dtr + - (r0_6*X[6])*P46*w0_6;
        const float f16 = X[16] + ( - (r1_6*X[16]) + (r0_6*X[6]))*dtr + + (r0_6*X[6])*P46*w0_6;
        const float f26 = X[26] + ( - (r2_6*X[26]) + (r1_6*X[16]))*dtr;
        s36 += ( + (r1_6*X[16]))*dtr;
        X[6] = plom_clipf(f6);
        X[16] = plom_clipf(f16);
        const float r0_7 = plom_fdivf((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])), C7);
        const float r1_7 = P17;
        const float r2_7 = P27;
        const float r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        const float w0_7 = sqdt * zb3;
        const float f7 = X[7] + ( - (r0_7*X[7]) + (r3_7))*dtr + - (r0_7*X[7])*P47*w0_7;
        const float f17 = X[17] + ( - (r1_7*X[17]) + (r0_7*X[7]))*dtr + + (r0_7*X[7])*P47*w0_7;
        const float f27 = X[27] + ( - (r2_7*X[27]) + (r1_7*X[17]))*dtr;
        s37 += ( + (r1_7*X[17]))*dtr;
        X[7] = plom_clipf(f7);
        X[17] = plom_clipf(f17);
        const float r0_8 = plom_fdivf((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])), C8);
        const float r1_8 = P18;
        const float r2_8 = P28;
        const float r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        float za4, zb4;
        plom_normal_pair32(ns, &za4, &zb4);
        const float w0_8 = sqdt * za4;
        const float f8 = X[8] + ( - (r0_8*X[8]) + (r3_8))*dtr + - (r0_8*X[8])*P48*w0_8;
        const float f18 = X[18] + ( - (r1_8*X[18]) + (r0_8*X[8]))*dtr + + (r0_8*X[8])*P48*w0_8;
        const float f28 = X[28] + ( - (r2_8*X[28]) + (r1_8*X[18]))*dtr;
        s38 += ( + (r1_8*X[18]))*dtr;
        X[8] = plom_clipf(f8);
        X[18] = plom_clipf(f18);
        const float r0_9 = plom_fdivf((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])), C9);
        const float r1_9 = P19;
        const float r2_9 = P29;
        const float r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        const float w0_9 = sqdt * zb4;
        const float f9 = X[9] + ( - (r0_9*X[9]) + (r3_9))*dtr + - (r0_9*X[9])*P49*w0_9;
        const float f19 = X[19] + ( - (r1_9*X[19]) + (r0_9*X[9]))*dtr + + (r0_9*X[9])*P49*w0_9;
        const float f29 = X[29] + ( - (r2_9*X[29]) + (r1_9*X[19]))*dtr;
        s39 += ( + (r1_9*X[19]))*dtr;
        X[9] = plom_clipf(f9);
        X[19] = plom_clipf(f19);
        X[20] = plom_clipf(f20);
        X[21] = plom_clipf(f21);
        X[22] = plom_clipf(f22);
        X[23] = plom_clipf(f23);
        X[24] = plom_clipf(f24);
        X[25] = plom_clipf(f25);
        X[26] = plom_clipf(f26);
        X[27] = plom_clipf(f27);
        X[28] = plom_clipf(f28);
        X[29] = plom_clipf(f29);
        const float f30 = X[30] + s30;
        X[30] = plom_clipf(f30);
        const float f31 = X[31] + s31;
        X[31] = plom_clipf(f31);
        const float f32 = X[32] + s32;
        X[32] = plom_clipf(f32);
        const float f33 = X[33] + s33;
        X[33] = plom_clipf(f33);
        const float f34 = X[34] + s34;
        X[34] = plom_clipf(f34);
        const float f35 = X[35] + s35;
        X[35] = plom_clipf(f35);
        const float f36 = X[36] + s36;
        X[36] = plom_clipf(f36);
        const float f37 = X[37] + s37;
        X[37] = plom_clipf(f37);
        const float f38 = X[38] + s38;
        X[38] = plom_clipf(f38);
        const float f39 = X[39] + s39;
        X[39] = plom_clipf(f39);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_interval_no_env_sto_pp_f32(float *X, const float *pv, const float *lp, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
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
    const float P20 = lp[20];
    const float P21 = lp[21];
    const float P22 = lp[22];
    const float P23 = lp[23];
    const float P24 = lp[24];
    const float P25 = lp[25];
    const float P26 = lp[26];
    const float P27 = lp[27];
    const float P28 = lp[28];
    const float P29 = lp[29];
    const float P30 = lp[30];
    const float P31 = lp[31];
    const float P32 = lp[32];
    const float P33 = lp[33];
    const float P34 = lp[34];
    const float P35 = lp[35];
    const float P36 = lp[36];
    const float P37 = lp[37];
    const float P38 = lp[38];
    const float P39 = lp[39];
    const float P80 = PLOM_UNIFORM32(pv[80]);
    const float P81 = PLOM_UNIFORM32(pv[81]);
    const float P82 = PLOM_UNIFORM32(pv[82]);
    const float P83 = PLOM_UNIFORM32(pv[83]);
    const float P84 = PLOM_UNIFORM32(pv[84]);
    const float P85 = PLOM_UNIFORM32(pv[85]);
    const float P86 = PLOM_UNIFORM32(pv[86]);
    const float P87 = PLOM_UNIFORM32(pv[87]);
    const float P88 = PLOM_UNIFORM32(pv[88]);
    const float P89 = PLOM_UNIFORM32(pv[89]);
    const float P90 = PLOM_UNIFORM32(pv[90]);
    const float P91 = PLOM_UNIFORM32(pv[91]);
    const float P92 = PLOM_UNIFORM32(pv[92]);
    const float P93 = PLOM_UNIFORM32(pv[93]);
    const float P94 = PLOM_UNIFORM32(pv[94]);
    const float P95 = PLOM_UNIFORM32(pv[95]);
    const float P96 = PLOM_UNIFORM32(pv[96]);
    const float P97 = PLOM_UNIFORM32(pv[97]);
    const float P98 = PLOM_UNIFORM32(pv[98]);
    const float P99 = PLOM_UNIFORM32(pv[99]);
    const float P100 = PLOM_UNIFORM32(pv[100]);
    const float P101 = PLOM_UNIFORM32(pv[101]);
    const float P102 = PLOM_UNIFORM32(pv[102]);
    const float P103 = PLOM_UNIFORM32(pv[103]);
    const float P104 = PLOM_UNIFORM32(pv[104]);
    const float P105 = PLOM_UNIFORM32(pv[105]);
    const float P106 = PLOM_UNIFORM32(pv[106]);
    const float P107 = PLOM_UNIFORM32(pv[107]);
    const float P108 = PLOM_UNIFORM32(pv[108]);
    const float P109 = PLOM_UNIFORM32(pv[109]);
    const float P110 = PLOM_UNIFORM32(pv[110]);
    const float P111 = PLOM_UNIFORM32(pv[111]);
    const float P112 = PLOM_UNIFORM32(pv[112]);
    const float P113 = PLOM_UNIFORM32(pv[113]);
    const float P114 = PLOM_UNIFORM32(pv[114]);
    const float P115 = PLOM_UNIFORM32(pv[115]);
    const float P116 = PLOM_UNIFORM32(pv[116]);
    const float P117 = PLOM_UNIFORM32(pv[117]);
    const float P118 = PLOM_UNIFORM32(pv[118]);
    const float P119 = PLOM_UNIFORM32(pv[119]);
    const float P120 = PLOM_UNIFORM32(pv[120]);
    const float P121 = PLOM_UNIFORM32(pv[121]);
    const float P122 = PLOM_UNIFORM32(pv[122]);
    const float P123 = PLOM_UNIFORM32(pv[123]);
    const float P124 = PLOM_UNIFORM32(pv[124]);
    const float P125 = PLOM_UNIFORM32(pv[125]);
    const float P126 = PLOM_UNIFORM32(pv[126]);
    const float P127 = PLOM_UNIFORM32(pv[127]);
    const float P128 = PLOM_UNIFORM32(pv[128]);
    const float P129 = PLOM_UNIFORM32(pv[129]);
    const float P130 = PLOM_UNIFORM32(pv[130]);
    const float P131 = PLOM_UNIFORM32(pv[131]);
    const float P132 = PLOM_UNIFORM32(pv[132]);
    const float P133 = PLOM_UNIFORM32(pv[133]);
    const float P134 = PLOM_UNIFORM32(pv[134]);
    const float P135 = PLOM_UNIFORM32(pv[135]);
    const float P136 = PLOM_UNIFORM32(pv[136]);
    const float P137 = PLOM_UNIFORM32(pv[137]);
    const float P138 = PLOM_UNIFORM32(pv[138]);
    const float P139 = PLOM_UNIFORM32(pv[139]);
    const float P140 = PLOM_UNIFORM32(pv[140]);
    const float P141 = PLOM_UNIFORM32(pv[141]);
    const float P142 = PLOM_UNIFORM32(pv[142]);
    const float P143 = PLOM_UNIFORM32(pv[143]);
    const float P144 = PLOM_UNIFORM32(pv[144]);
    const float P145 = PLOM_UNIFORM32(pv[145]);
    const float P146 = PLOM_UNIFORM32(pv[146]);
    const float P147 = PLOM_UNIFORM32(pv[147]);
    const float P148 = PLOM_UNIFORM32(pv[148]);
    const float P149 = PLOM_UNIFORM32(pv[149]);
    const float P150 = PLOM_UNIFORM32(pv[150]);
    const float P151 = PLOM_UNIFORM32(pv[151]);
    const float P152 = PLOM_UNIFORM32(pv[152]);
    const float P153 = PLOM_UNIFORM32(pv[153]);
    const float P154 = PLOM_UNIFORM32(pv[154]);
    const float P155 = PLOM_UNIFORM32(pv[155]);
    const float P156 = PLOM_UNIFORM32(pv[156]);
    const float P157 = PLOM_UNIFORM32(pv[157]);
    const float P158 = PLOM_UNIFORM32(pv[158]);
    const float P159 = PLOM_UNIFORM32(pv[159]);
    const float P160 = PLOM_UNIFORM32(pv[160]);
    const float P161 = PLOM_UNIFORM32(pv[161]);
    const float P162 = PLOM_UNIFORM32(pv[162]);
    const float P163 = PLOM_UNIFORM32(pv[163]);
    const float P164 = PLOM_UNIFORM32(pv[164]);
    const float P165 = PLOM_UNIFORM32(pv[165]);
    const float P166 = PLOM_UNIFORM32(pv[166]);
    const float P167 = PLOM_UNIFORM32(pv[167]);
    const float P168 = PLOM_UNIFORM32(pv[168]);
    const float P169 = PLOM_UNIFORM32(pv[169]);
    const float P170 = PLOM_UNIFORM32(pv[170]);
    const float P171 = PLOM_UNIFORM32(pv[171]);
    const float P172 = PLOM_UNIFORM32(pv[172]);
    const float P173 = PLOM_UNIFORM32(pv[173]);
    const float P174 = PLOM_UNIFORM32(pv[174]);
    const float P175 = PLOM_UNIFORM32(pv[175]);
    const float P176 = PLOM_UNIFORM32(pv[176]);
    const float P177 = PLOM_UNIFORM32(pv[177]);
    const float P178 = PLOM_UNIFORM32(pv[178]);
    const float P179 = PLOM_UNIFORM32(pv[179]);
    X[30] = 0.0f; /* reset_inc */
    X[31] = 0.0f; /* reset_inc */
    X[32] = 0.0f; /* reset_inc */
    X[33] = 0.0f; /* reset_inc */
    X[34] = 0.0f; /* reset_inc */
    X[35] = 0.0f; /* reset_inc */
    X[36] = 0.0f; /* reset_inc */
    X[37] = 0.0f; /* reset_inc */
    X[38] = 0.0f; /* reset_inc */
    X[39] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        const float C1 = (float)plom_cov_ld(cvk + 1);
        const float C2 = (float)plom_cov_ld(cvk + 2);
        const float C3 = (float)plom_cov_ld(cvk + 3);
        const float C4 = (float)plom_cov_ld(cvk + 4);
        const float C5 = (float)plom_cov_ld(cvk + 5);
        const float C6 = (float)plom_cov_ld(cvk + 6);
        const float C7 = (float)plom_cov_ld(cvk + 7);
        const float C8 = (float)plom_cov_ld(cvk + 8);
        const float C9 = (float)plom_cov_ld(cvk + 9);
        float s30 = 0.0f;
        float s31 = 0.0f;
        float s32 = 0.0f;
        float s33 = 0.0f;
        float s34 = 0.0f;
        float s35 = 0.0f;
        float s36 = 0.0f;
        float s37 = 0.0f;
        float s38 = 0.0f;
        float s39 = 0.0f;
        const float r0_0 = plom_fdivf((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])), C0);
        const float r1_0 = P10;
        const float r2_0 = P20;
        const float r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        float za0, zb0;
        plom_normal_pair32(ns, &za0, &zb0);
        const float d0_0 = sqdt * za0;
        const float d1_0 = sqdt * zb0;
        float za1, zb1;
        plom_normal_pair32(ns, &za1, &zb1);
        const float d2_0 = sqdt * za1;
        const float d3_0 = sqdt * zb1;
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r3_0))*dtr + - plom_fsqrtf((r0_0*X[0]))*d0_0+ plom_fsqrtf((r3_0))*d3_0;
        const float f10 = X[10] + ( - (r1_0*X[10]) + (r0_0*X[0]))*dtr + - plom_fsqrtf((r1_0*X[10]))*d1_0+ plom_fsqrtf((r0_0*X[0]))*d0_0;
        const float f20 = X[20] + ( - (r2_0*X[20]) + (r1_0*X[10]))*dtr + - plom_fsqrtf((r2_0*X[20]))*d2_0+ plom_fsqrtf((r1_0*X[10]))*d1_0;
        s30 += ( + (r1_0*X[10]))*dtr + + plom_fsqrtf((r1_0*X[10]))*d1_0;
        X[0] = plom_clipf(f0);
        X[10] = plom_clipf(f10);
        const float r0_1 = plom_fdivf((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])), C1);
        const float r1_1 = P11;
        const float r2_1 = P21;
        const float r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        float za2, zb2;
        plom_normal_pair32(ns, &za2, &zb2);
        const float d0_1 = sqdt * za2;
        const float d1_1 = sqdt * zb2;
        float za3, zb3;
        plom_normal_pair32(ns, &za3, &zb3);
        const float d2_1 = sqdt * za3;
        const float d3_1 = sqdt * zb3;
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r3_1))*dtr + - plom_fsqrtf((r0_1*X[1]))*d0_1+ plom_fsqrtf((r3_1))*d3_1;
        const float f11 = X[11] + ( - (r1_1*X[11]) + (r0_1*X[1]))*dtr + - plom_fsqrtf((r1_1*X[11]))*d1_1+ plom_fsqrtf((r0_1*X[1]))*d0_1;
        const float f21 = X[21] + ( - (r2_1*X[21]) + (r1_1*X[11]))*dtr + - plom_fsqrtf((r2_1*X[21]))*d2_1+ plom_fsqrtf((r1_1*X[11]))*d1_1;
        s31 += ( + (r1_1*X[11]))*dtr + + plom_fsqrtf((r1_1*X[11]))*d1_1;
        X[1] = plom_clipf(f1);
        X[11] = plom_clipf(f11);
        const float r0_2 = plom_fdivf((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])), C2);
        const float r1_2 = P12;
        const float r2_2 = P22;
        const float r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        float za4, zb4;
        plom_normal_pair32(ns, &za4, &zb4);
        const float d0_2 = sqdt * za4;
        const float d1_2 = sqdt * zb4;
        float za5, zb5;
        plom_normal_pair32(ns, &za5, &zb5);
        const float d2_2 = sqdt * za5;
        const float d3_2 = sqdt * zb5;
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r3_2))*dtr + - plom_fsqrtf((r0_2*X[2]))*d0_2+ plom_fsqrtf((r3_2))*d3_2;
        const float f12 = X[12] + ( - (r1_2*X[12]) + (r0_2*X[2]))*dtr + - plom_fsqrtf((r1_2*X[12]))*d1_2+ plom_fsqrtf((r0_2*X[2]))*d0_2;
        const float f22 = X[22] + ( - (r2_2*X[22]) + (r1_2*X[12]))*dtr + - plom_fsqrtf((r2_2*X[22]))*d2_2+ plom_fsqrtf((r1_2*X[12]))*d1_2;
        s32 += ( + (r1_2*X[12]))*dtr + + plom_fsqrtf((r1_2*X[12]))*d1_2;
        X[2] = plom_clipf(f2);
        X[12] = plom_clipf(f12);
        const float r0_3 = plom_fdivf((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])), C3);
        const float r1_3 = P13;
        const float r2_3 = P23;
        const float r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        float za6, zb6;
        plom_normal_pair32(ns, &za6, &zb6);
        const float d0_3 = sqdt * za6;
        const float d1_3 = sqdt * zb6;
        float za7, zb7;
        plom_normal_pair32(ns, &za7, &zb7);
        const float d2_3 = sqdt * za7;
        const float d3_3 = sqdt * zb7;
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r3_3))*dtr + - plom_fsqrtf((r0_3*X[3]))*d0_3+ plom_fsqrtf((r3_3))*d3_3;
        const float f13 = X[13] + ( - (r1_3*X[13]) + (r0_3*X[3]))*dtr + - plom_fsqrtf((r1_3*X[13]))*d1_3+ plom_fsqrtf((r0_3*X[3]))*d0_3;
        const float f23 = X[23] + ( - (r2_3*X[23]) + (r1_3*X[13]))*dtr + - plom_fsqrtf((r2_3*X[23]))*d2_3+ plom_fsqrtf((r1_3*X[13]))*d1_3;
        s33 += ( + (r1_3*X[13]))*dtr + + plom_fsqrtf((r1_3*X[13]))*d1_3;
        X[3] = plom_clipf(f3);
        X[13] = plom_clipf(f13);
        const float r0_4 = plom_fdivf((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])), C4);
        const float r1_4 = P14;
        const float r2_4 = P24;
        const float r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        float za8, zb8;
        plom_normal_pair32(ns, &za8, &zb8);
        const float d0_4 = sqdt * za8;
        const float d1_4 = sqdt * zb8;
        float za9, zb9;
        plom_normal_pair32(ns, &za9, &zb9);
        const float d2_4 = sqdt * za9;
        const float d3_4 = sqdt * zb9;
        const float f4 = X[4] + ( - (r0_4*X[4]) + (r3_4))*dtr + - plom_fsqrtf((r0_4*X[4]))*d0_4+ plom_fsqrtf((r3_4))*d3_4;
        const float f14 = X[14] + ( - (r1_4*X[14]) + (r0_4*X[4]))*dtr + - plom_fsqrtf((r1_4*X[14]))*d1_4+ plom_fsqrtf((r0_4*X[4]))*d0_4;
        const float f24 = X[24] + ( - (r2_4*X[24]) + (r1_4*X[14]))*dtr + - plom_fsqrtf((r2_4*X[24]))*d2_4+ plom_fsqrtf((r1_4*X[14]))*d1_4;
        s34 += ( + (r1_4*X[14]))*dtr + + plom_fsqrtf((r1_4*X[14]))*d1_4;
        X[4] = plom_clipf(f4);
        X[14] = plom_clipf(f14);
        const float r0_5 = plom_fdivf((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])), C5);
        const float r1_5 = P15;
        const float r2_5 = P25;
        const float r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        float za10, zb10;
        plom_normal_pair32(ns, &za10, &zb10);
        const float d0_5 = sqdt * za10;
        const float d1_5 = sqdt * zb10;
        float za11, zb11;
        plom_normal_pair32(ns, &za11, &zb11);
        const float d2_5 = sqdt * za11;
        const float d3_5 = sqdt * zb11;
        const float f5 = X[5] + ( - (r0_5*X[5]) + (r3_5))*dtr + - plom_fsqrtf((r0_5*X[5]))*d0_5+ plom_fsqrtf((r3_5))*d3_5;
        const float f15 = X[15] + ( - (r1_5*X[15]) + (r0_5*X[5]))*dtr + - plom_fsqrtf((r1_5*X[15]))*d1_5+ plom_fsqrtf((r0_5*X[5]))*d0_5;
        const float f25 = X[25] + ( - (r2_5*X[25]) + (r1_5*X[15]))*dtr + - plom_fsqrtf((r2_5*X[25]))*d2_5+ plom_fsqrtf((r1_5*X[15]))*d1_5;
        s35 += ( + (r1_5*X[15]))*dtr + + plom_fsqrtf((r1_5*X[15]))*d1_5;
        X[5] = plom_clipf(f5);
        X[15] = plom_clipf(f15);
        const float r0_6 = plom_fdivf((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])), C6);
        const float r1_6 = P16;
        const float r2_6 = P26;
        const float r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        float za12, zb12;
        plom_normal_pair32(ns, &za12, &zb12);
        const float d0_6 = sqdt * za12;
        const float d1_6 = sqdt * zb12;
        float za13, zb13;
        plom_normal_pair32(ns, &za13, &zb13);
        const float d2_6 = sqdt * za13;
        const float d3_6 = sqdt * zb13;
        const float f6 = X[6] + ( - (r0_6*X[6]) + (r3_6))*dtr + - plom_fsqrtf((r0_6*X[6]))*d0_6+ plom_fsqrtf((r3_6))*d3_6;
        const float f16 = X[16] + ( - (r1_6*X[16]) + (r0_6*X[6]))*dtr + - plom_fsqrtf((r1_6*X[16]))*d1_6+ plom_fsqrtf((r0_6*X[6]))*d0_6;
        const float f26 = X[26] + ( - (r2_6*X[26]) + (r1_6*X[16]))*dtr + - plom_fsqrtf((r2_6*X[26]))*d2_6+ plom_fsqrtf((r1_6*X[16]))*d1_6;
        s36 += ( + (r1_6*X[16]))*dtr + + plom_fsqrtf((r1_6*X[16]))*d1_6;
        X[6] = plom_clipf(f6);
        X[16] = plom_clipf(f16);
        const float r0_7 = plom_fdivf((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])), C7);
        const float r1_7 = P17;
        const float r2_7 = P27;
        const float r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        float za14, zb14;
        plom_normal_pair32(ns, &za14, &zb14);
        const float d0_7 = sqdt * za14;
        const float d1_7 = sqdt * zb14;
        float za15, zb15;
        plom_normal_pair32(ns, &za15, &zb15);
        const float d2_7 = sqdt * za15;
        const float d3_7 = sqdt * zb15;
        const float f7 = X[7] + ( - (r0_7*X[7]) + (r3_7))*dtr + - plom_fsqrtf((r0_7*X[7]))*d0_7+ plom_fsqrtf((r3_7))*d3_7;
        const float f17 = X[17] + ( - (r1_7*X[17]) + (r0_7*X[7]))*dtr + - plom_fsqrtf((r1_7*X[17]))*d1_7+ plom_fsqrtf((r0_7*X[7]))*d0_7;
        const float f27 = X[27] + ( - (r2_7*X[27]) + (r1_7*X[17]))*dtr + - plom_fsqrtf((r2_7*X[27]))*d2_7+ plom_fsqrtf((r1_7*X[17]))*d1_7;
        s37 += ( + (r1_7*X[17]))*dtr + + plom_fsqrtf((r1_7*X[17]))*d1_7;
        X[7] = plom_clipf(f7);
        X[17] = plom_clipf(f17);
        const float r0_8 = plom_fdivf((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])), C8);
        const float r1_8 = P18;
        const float r2_8 = P28;
        const float r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        float za16, zb16;
        plom_normal_pair32(ns, &za16, &zb16);
        const float d0_8 = sqdt * za16;
        const float d1_8 = sqdt * zb16;
        float za17, zb17;
        plom_normal_pair32(ns, &za17, &zb17);
        const float d2_8 = sqdt * za17;
        const float d3_8 = sqdt * zb17;
        const float f8 = X[8] + ( - (r0_8*X[8]) + (r3_8))*dtr + - plom_fsqrtf((r0_8*X[8]))*d0_8+ plom_fsqrtf((r3_8))*d3_8;
        const float f18 = X[18] + ( - (r1_8*X[18]) + (r0_8*X[8]))*dtr + - plom_fsqrtf((r1_8*X[18]))*d1_8+ plom_fsqrtf((r0_8*X[8]))*d0_8;
        const float f28 = X[28] + ( - (r2_8*X[28]) + (r1_8*X[18]))*dtr + - plom_fsqrtf((r2_8*X[28]))*d2_8+ plom_fsqrtf((r1_8*X[18]))*d1_8;
        s38 += ( + (r1_8*X[18]))*dtr + + plom_fsqrtf((r1_8*X[18]))*d1_8;
        X[8] = plom_clipf(f8);
        X[18] = plom_clipf(f18);
        const float r0_9 = plom_fdivf((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])), C9);
        const float r1_9 = P19;
        const float r2_9 = P29;
        const float r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        float za18, zb18;
        plom_normal_pair32(ns, &za18, &zb18);
        const float d0_9 = sqdt * za18;
        const float d1_9 = sqdt * zb18;
        float za19, zb19;
        plom_normal_pair32(ns, &za19, &zb19);
        const float d2_9 = sqdt * za19;
        const float d3_9 = sqdt * zb19;
        const float f9 = X[9] + ( - (r0_9*X[9]) + (r3_9))*dtr + - plom_fsqrtf((r0_9*X[9]))*d0_9+ plom_fsqrtf((r3_9))*d3_9;
        const float f19 = X[19] + ( - (r1_9*X[19]) + (r0_9*X[9]))*dtr + - plom_fsqrtf((r1_9*X[19]))*d1_9+ plom_fsqrtf((r0_9*X[9]))*d0_9;
        const float f29 = X[29] + ( - (r2_9*X[29]) + (r1_9*X[19]))*dtr + - plom_fsqrtf((r2_9*X[29]))*d2_9+ plom_fsqrtf((r1_9*X[19]))*d1_9;
        s39 += ( + (r1_9*X[19]))*dtr + + plom_fsqrtf((r1_9*X[19]))*d1_9;
        X[9] = plom_clipf(f9);
        X[19] = plom_clipf(f19);
        X[20] = plom_clipf(f20);
        X[21] = plom_clipf(f21);
        X[22] = plom_clipf(f22);
        X[23] = plom_clipf(f23);
        X[24] = plom_clipf(f24);
        X[25] = plom_clipf(f25);
        X[26] = plom_clipf(f26);
        X[27] = plom_clipf(f27);
        X[28] = plom_clipf(f28);
        X[29] = plom_clipf(f29);
        const float f30 = X[30] + s30;
        X[30] = plom_clipf(f30);
        const float f31 = X[31] + s31;
        X[31] = plom_clipf(f31);
        const float f32 = X[32] + s32;
        X[32] = plom_clipf(f32);
        const float f33 = X[33] + s33;
        X[33] = plom_clipf(f33);
        const float f34 = X[34] + s34;
        X[34] = plom_clipf(f34);
        const float f35 = X[35] + s35;
        X[35] = plom_clipf(f35);
        const float f36 = X[36] + s36;
        X[36] = plom_clipf(f36);
        const float f37 = X[37] + s37;
        X[37] = plom_clipf(f37);
        const float f38 = X[38] + s38;
        X[38] = plom_clipf(f38);
        const float f39 = X[39] + s39;
        X[39] = plom_clipf(f39);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_interval_no_dem_sto_no_env_sto_pp_f32(float *X, const float *pv, const float *lp, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
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
    const float P20 = lp[20];
    const float P21 = lp[21];
    const float P22 = lp[22];
    const float P23 = lp[23];
    const float P24 = lp[24];
    const float P25 = lp[25];
    const float P26 = lp[26];
    const float P27 = lp[27];
    const float P28 = lp[28];
    const float P29 = lp[29];
    const float P30 = lp[30];
    const float P31 = lp[31];
    const float P32 = lp[32];
    const float P33 = lp[33];
    const float P34 = lp[34];
    const float P35 = lp[35];
    const float P36 = lp[36];
    const float P37 = lp[37];
    const float P38 = lp[38];
    const float P39 = lp[39];
    const float P80 = PLOM_UNIFORM32(pv[80]);
    const float P81 = PLOM_UNIFORM32(pv[81]);
    const float P82 = PLOM_UNIFORM32(pv[82]);
    const float P83 = PLOM_UNIFORM32(pv[83]);
    const float P84 = PLOM_UNIFORM32(pv[84]);
    const float P85 = PLOM_UNIFORM32(pv[85]);
    const float P86 = PLOM_UNIFORM32(pv[86]);
    const float P87 = PLOM_UNIFORM32(pv[87]);
    const float P88 = PLOM_UNIFORM32(pv[88]);
    const float P89 = PLOM_UNIFORM32(pv[89]);
    const float P90 = PLOM_UNIFORM32(pv[90]);
    const float P91 = PLOM_UNIFORM32(pv[91]);
    const float P92 = PLOM_UNIFORM32(pv[92]);
    const float P93 = PLOM_UNIFORM32(pv[93]);
    const float P94 = PLOM_UNIFORM32(pv[94]);
    const float P95 = PLOM_UNIFORM32(pv[95]);
    const float P96 = PLOM_UNIFORM32(pv[96]);
    const float P97 = PLOM_UNIFORM32(pv[97]);
    const float P98 = PLOM_UNIFORM32(pv[98]);
    const float P99 = PLOM_UNIFORM32(pv[99]);
    const float P100 = PLOM_UNIFORM32(pv[100]);
    const float P101 = PLOM_UNIFORM32(pv[101]);
    const float P102 = PLOM_UNIFORM32(pv[102]);
    const float P103 = PLOM_UNIFORM32(pv[103]);
    const float P104 = PLOM_UNIFORM32(pv[104]);
    const float P105 = PLOM_UNIFORM32(pv[105]);
    const float P106 = PLOM_UNIFORM32(pv[106]);
    const float P107 = PLOM_UNIFORM32(pv[107]);
    const float P108 = PLOM_UNIFORM32(pv[108]);
    const float P109 = PLOM_UNIFORM32(pv[109]);
    const float P110 = PLOM_UNIFORM32(pv[110]);
    const float P111 = PLOM_UNIFORM32(pv[111]);
    const float P112 = PLOM_UNIFORM32(pv[112]);
    const float P113 = PLOM_UNIFORM32(pv[113]);
    const float P114 = PLOM_UNIFORM32(pv[114]);
    const float P115 = PLOM_UNIFORM32(pv[115]);
    const float P116 = PLOM_UNIFORM32(pv[116]);
    const float P117 = PLOM_UNIFORM32(pv[117]);
    const float P118 = PLOM_UNIFORM32(pv[118]);
    const float P119 = PLOM_UNIFORM32(pv[119]);
    const float P120 = PLOM_UNIFORM32(pv[120]);
    const float P121 = PLOM_UNIFORM32(pv[121]);
    const float P122 = PLOM_UNIFORM32(pv[122]);
    const float P123 = PLOM_UNIFORM32(pv[123]);
    const float P124 = PLOM_UNIFORM32(pv[124]);
    const float P125 = PLOM_UNIFORM32(pv[125]);
    const float P126 = PLOM_UNIFORM32(pv[126]);
    const float P127 = PLOM_UNIFORM32(pv[127]);
    const float P128 = PLOM_UNIFORM32(pv[128]);
    const float P129 = PLOM_UNIFORM32(pv[129]);
    const float P130 = PLOM_UNIFORM32(pv[130]);
    const float P131 = PLOM_UNIFORM32(pv[131]);
    const float P132 = PLOM_UNIFORM32(pv[132]);
    const float P133 = PLOM_UNIFORM32(pv[133]);
    const float P134 = PLOM_UNIFORM32(pv[134]);
    const float P135 = PLOM_UNIFORM32(pv[135]);
    const float P136 = PLOM_UNIFORM32(pv[136]);
    const float P137 = PLOM_UNIFORM32(pv[137]);
    const float P138 = PLOM_UNIFORM32(pv[138]);
    const float P139 = PLOM_UNIFORM32(pv[139]);
    const float P140 = PLOM_UNIFORM32(pv[140]);
    const float P141 = PLOM_UNIFORM32(pv[141]);
    const float P142 = PLOM_UNIFORM32(pv[142]);
    const float P143 = PLOM_UNIFORM32(pv[143]);
    const float P144 = PLOM_UNIFORM32(pv[144]);
    const float P145 = PLOM_UNIFORM32(pv[145]);
    const float P146 = PLOM_UNIFORM32(pv[146]);
    const float P147 = PLOM_UNIFORM32(pv[147]);
    const float P148 = PLOM_UNIFORM32(pv[148]);
    const float P149 = PLOM_UNIFORM32(pv[149]);
    const float P150 = PLOM_UNIFORM32(pv[150]);
    const float P151 = PLOM_UNIFORM32(pv[151]);
    const float P152 = PLOM_UNIFORM32(pv[152]);
    const float P153 = PLOM_UNIFORM32(pv[153]);
    const float P154 = PLOM_UNIFORM32(pv[154]);
    const float P155 = PLOM_UNIFORM32(pv[155]);
    const float P156 = PLOM_UNIFORM32(pv[156]);
    const float P157 = PLOM_UNIFORM32(pv[157]);
    const float P158 = PLOM_UNIFORM32(pv[158]);
    const float P159 = PLOM_UNIFORM32(pv[159]);
    const float P160 = PLOM_UNIFORM32(pv[160]);
    const float P161 = PLOM_UNIFORM32(pv[161]);
    const float P162 = PLOM_UNIFORM32(pv[162]);
    const float P163 = PLOM_UNIFORM32(pv[163]);
    const float P164 = PLOM_UNIFORM32(pv[164]);
    const float P165 = PLOM_UNIFORM32(pv[165]);
    const float P166 = PLOM_UNIFORM32(pv[166]);
    const float P167 = PLOM_UNIFORM32(pv[167]);
    const float P168 = PLOM_UNIFORM32(pv[168]);
    const float P169 = PLOM_UNIFORM32(pv[169]);
    const float P170 = PLOM_UNIFORM32(pv[170]);
    const float P171 = PLOM_UNIFORM32(pv[171]);
    const float P172 = PLOM_UNIFORM32(pv[172]);
    const float P173 = PLOM_UNIFORM32(pv[173]);
    const float P174 = PLOM_UNIFORM32(pv[174]);
    const float P175 = PLOM_UNIFORM32(pv[175]);
    const float P176 = PLOM_UNIFORM32(pv[176]);
    const float P177 = PLOM_UNIFORM32(pv[177]);
    const float P178 = PLOM_UNIFORM32(pv[178]);
    const float P179 = PLOM_UNIFORM32(pv[179]);
    X[30] = 0.0f; /* reset_inc */
    X[31] = 0.0f; /* reset_inc */
    X[32] = 0.0f; /* reset_inc */
    X[33] = 0.0f; /* reset_inc */
    X[34] = 0.0f; /* reset_inc */
    X[35] = 0.0f; /* reset_inc */
    X[36] = 0.0f; /* reset_inc */
    X[37] = 0.0f; /* reset_inc */
    X[38] = 0.0f; /* reset_inc */
    X[39] = 0.0f; /* reset_inc */
    double t = t0;
    const int K = PLOM_UNIFORM_INT(plom_substeps(t0, t1, dt));
    for (int k = 0; k < K; k++) {
        const float tr = (float)t; (void)tr;
        const double *cvk = cv + (size_t)k * (size_t)cs;
        const float C0 = (float)plom_cov_ld(cvk + 0);
        const float C1 = (float)plom_cov_ld(cvk + 1);
        const float C2 = (float)plom_cov_ld(cvk + 2);
        const float C3 = (float)plom_cov_ld(cvk + 3);
        const float C4 = (float)plom_cov_ld(cvk + 4);
        const float C5 = (float)plom_cov_ld(cvk + 5);
        const float C6 = (float)plom_cov_ld(cvk + 6);
        const float C7 = (float)plom_cov_ld(cvk + 7);
        const float C8 = (float)plom_cov_ld(cvk + 8);
        const float C9 = (float)plom_cov_ld(cvk + 9);
        float s30 = 0.0f;
        float s31 = 0.0f;
        float s32 = 0.0f;
        float s33 = 0.0f;
        float s34 = 0.0f;
        float s35 = 0.0f;
        float s36 = 0.0f;
        float s37 = 0.0f;
        float s38 = 0.0f;
        float s39 = 0.0f;
        const float r0_0 = plom_fdivf((P0 * (P80 * X[20] + P81 * X[21] + P82 * X[22] + P83 * X[23] + P84 * X[24] + P85 * X[25] + P86 * X[26] + P87 * X[27] + P88 * X[28] + P89 * X[29])), C0);
        const float r1_0 = P10;
        const float r2_0 = P20;
        const float r3_0 = (P30 * (((C0 - X[0]) - X[10]) - X[20]));
        const float f0 = X[0] + ( - (r0_0*X[0]) + (r3_0))*dtr;
        const float f10 = X[10] + ( - (r1_0*X[10]) + (r0_0*X[0]))*dtr;
        const float f20 = X[20] + ( - (r2_0*X[20]) + (r1_0*X[10]))*dtr;
        s30 += ( + (r1_0*X[10]))*dtr;
        X[0] = plom_clipf(f0);
        X[10] = plom_clipf(f10);
        const float r0_1 = plom_fdivf((P1 * (P90 * X[20] + P91 * X[21] + P92 * X[22] + P93 * X[23] + P94 * X[24] + P95 * X[25] + P96 * X[26] + P97 * X[27] + P98 * X[28] + P99 * X[29])), C1);
        const float r1_1 = P11;
        const float r2_1 = P21;
        const float r3_1 = (P31 * (((C1 - X[1]) - X[11]) - X[21]));
        const float f1 = X[1] + ( - (r0_1*X[1]) + (r3_1))*dtr;
        const float f11 = X[11] + ( - (r1_1*X[11]) + (r0_1*X[1]))*dtr;
        const float f21 = X[21] + ( - (r2_1*X[21]) + (r1_1*X[11]))*dtr;
        s31 += ( + (r1_1*X[11]))*dtr;
        X[1] = plom_clipf(f1);
        X[11] = plom_clipf(f11);
        const float r0_2 = plom_fdivf((P2 * (P100 * X[20] + P101 * X[21] + P102 * X[22] + P103 * X[23] + P104 * X[24] + P105 * X[25] + P106 * X[26] + P107 * X[27] + P108 * X[28] + P109 * X[29])), C2);
        const float r1_2 = P12;
        const float r2_2 = P22;
        const float r3_2 = (P32 * (((C2 - X[2]) - X[12]) - X[22]));
        const float f2 = X[2] + ( - (r0_2*X[2]) + (r3_2))*dtr;
        const float f12 = X[12] + ( - (r1_2*X[12]) + (r0_2*X[2]))*dtr;
        const float f22 = X[22] + ( - (r2_2*X[22]) + (r1_2*X[12]))*dtr;
        s32 += ( + (r1_2*X[12]))*dtr;
        X[2] = plom_clipf(f2);
        X[12] = plom_clipf(f12);
        const float r0_3 = plom_fdivf((P3 * (P110 * X[20] + P111 * X[21] + P112 * X[22] + P113 * X[23] + P114 * X[24] + P115 * X[25] + P116 * X[26] + P117 * X[27] + P118 * X[28] + P119 * X[29])), C3);
        const float r1_3 = P13;
        const float r2_3 = P23;
        const float r3_3 = (P33 * (((C3 - X[3]) - X[13]) - X[23]));
        const float f3 = X[3] + ( - (r0_3*X[3]) + (r3_3))*dtr;
        const float f13 = X[13] + ( - (r1_3*X[13]) + (r0_3*X[3]))*dtr;
        const float f23 = X[23] + ( - (r2_3*X[23]) + (r1_3*X[13]))*dtr;
        s33 += ( + (r1_3*X[13]))*dtr;
        X[3] = plom_clipf(f3);
        X[13] = plom_clipf(f13);
        const float r0_4 = plom_fdivf((P4 * (P120 * X[20] + P121 * X[21] + P122 * X[22] + P123 * X[23] + P124 * X[24] + P125 * X[25] + P126 * X[26] + P127 * X[27] + P128 * X[28] + P129 * X[29])), C4);
        const float r1_4 = P14;
        const float r2_4 = P24;
        const float r3_4 = (P34 * (((C4 - X[4]) - X[14]) - X[24]));
        const float f4 = X[4] + ( - (r0_4*X[4]) + (r3_4))*dtr;
        const float f14 = X[14] + ( - (r1_4*X[14]) + (r0_4*X[4]))*dtr;
        const float f24 = X[24] + ( - (r2_4*X[24]) + (r1_4*X[14]))*dtr;
        s34 += ( + (r1_4*X[14]))*dtr;
        X[4] = plom_clipf(f4);
        X[14] = plom_clipf(f14);
        const float r0_5 = plom_fdivf((P5 * (P130 * X[20] + P131 * X[21] + P132 * X[22] + P133 * X[23] + P134 * X[24] + P135 * X[25] + P136 * X[26] + P137 * X[27] + P138 * X[28] + P139 * X[29])), C5);
        const float r1_5 = P15;
        const float r2_5 = P25;
        const float r3_5 = (P35 * (((C5 - X[5]) - X[15]) - X[25]));
        const float f5 = X[5] + ( - (r0_5*X[5]) + (r3_5))*dtr;
        const float f15 = X[15] + ( - (r1_5*X[15]) + (r0_5*X[5]))*dtr;
        const float f25 = X[25] + ( - (r2_5*X[25]) + (r1_5*X[15]))*dtr;
        s35 += ( + (r1_5*X[15]))*dtr;
        X[5] = plom_clipf(f5);
        X[15] = plom_clipf(f15);
        const float r0_6 = plom_fdivf((P6 * (P140 * X[20] + P141 * X[21] + P142 * X[22] + P143 * X[23] + P144 * X[24] + P145 * X[25] + P146 * X[26] + P147 * X[27] + P148 * X[28] + P149 * X[29])), C6);
        const float r1_6 = P16;
        const float r2_6 = P26;
        const float r3_6 = (P36 * (((C6 - X[6]) - X[16]) - X[26]));
        const float f6 = X[6] + ( - (r0_6*X[6]) + (r3_6))*dtr;
        const float f16 = X[16] + ( - (r1_6*X[16]) + (r0_6*X[6]))*dtr;
        const float f26 = X[26] + ( - (r2_6*X[26]) + (r1_6*X[16]))*dtr;
        s36 += ( + (r1_6*X[16]))*dtr;
        X[6] = plom_clipf(f6);
        X[16] = plom_clipf(f16);
        const float r0_7 = plom_fdivf((P7 * (P150 * X[20] + P151 * X[21] + P152 * X[22] + P153 * X[23] + P154 * X[24] + P155 * X[25] + P156 * X[26] + P157 * X[27] + P158 * X[28] + P159 * X[29])), C7);
        const float r1_7 = P17;
        const float r2_7 = P27;
        const float r3_7 = (P37 * (((C7 - X[7]) - X[17]) - X[27]));
        const float f7 = X[7] + ( - (r0_7*X[7]) + (r3_7))*dtr;
        const float f17 = X[17] + ( - (r1_7*X[17]) + (r0_7*X[7]))*dtr;
        const float f27 = X[27] + ( - (r2_7*X[27]) + (r1_7*X[17]))*dtr;
        s37 += ( + (r1_7*X[17]))*dtr;
        X[7] = plom_clipf(f7);
        X[17] = plom_clipf(f17);
        const float r0_8 = plom_fdivf((P8 * (P160 * X[20] + P161 * X[21] + P162 * X[22] + P163 * X[23] + P164 * X[24] + P165 * X[25] + P166 * X[26] + P167 * X[27] + P168 * X[28] + P169 * X[29])), C8);
        const float r1_8 = P18;
        const float r2_8 = P28;
        const float r3_8 = (P38 * (((C8 - X[8]) - X[18]) - X[28]));
        const float f8 = X[8] + ( - (r0_8*X[8]) + (r3_8))*dtr;
        const float f18 = X[18] + ( - (r1_8*X[18]) + (r0_8*X[8]))*dtr;
        const float f28 = X[28] + ( - (r2_8*X[28]) + (r1_8*X[18]))*dtr;
        s38 += ( + (r1_8*X[18]))*dtr;
        X[8] = plom_clipf(f8);
        X[18] = plom_clipf(f18);
        const float r0_9 = plom_fdivf((P9 * (P170 * X[20] + P171 * X[21] + P172 * X[22] + P173 * X[23] + P174 * X[24] + P175 * X[25] + P176 * X[26] + P177 * X[27] + P178 * X[28] + P179 * X[29])), C9);
        const float r1_9 = P19;
        const float r2_9 = P29;
        const float r3_9 = (P39 * (((C9 - X[9]) - X[19]) - X[29]));
        const float f9 = X[9] + ( - (r0_9*X[9]) + (r3_9))*dtr;
        const float f19 = X[19] + ( - (r1_9*X[19]) + (r0_9*X[9]))*dtr;
        const float f29 = X[29] + ( - (r2_9*X[29]) + (r1_9*X[19]))*dtr;
        s39 += ( + (r1_9*X[19]))*dtr;
        X[9] = plom_clipf(f9);
        X[19] = plom_clipf(f19);
        X[20] = plom_clipf(f20);
        X[21] = plom_clipf(f21);
        X[22] = plom_clipf(f22);
        X[23] = plom_clipf(f23);
        X[24] = plom_clipf(f24);
        X[25] = plom_clipf(f25);
        X[26] = plom_clipf(f26);
        X[27] = plom_clipf(f27);
        X[28] = plom_clipf(f28);
        X[29] = plom_clipf(f29);
        const float f30 = X[30] + s30;
        X[30] = plom_clipf(f30);
        const float f31 = X[31] + s31;
        X[31] = plom_clipf(f31);
        const float f32 = X[32] + s32;
        X[32] = plom_clipf(f32);
        const float f33 = X[33] + s33;
        X[33] = plom_clipf(f33);
        const float f34 = X[34] + s34;
        X[34] = plom_clipf(f34);
        const float f35 = X[35] + s35;
        X[35] = plom_clipf(f35);
        const float f36 = X[36] + s36;
        X[36] = plom_clipf(f36);
        const float f37 = X[37] + s37;
        X[37] = plom_clipf(f37);
        const float f38 = X[38] + s38;
        X[38] = plom_clipf(f38);
        const float f39 = X[39] + s39;
        X[39] = plom_clipf(f39);
        t += dt;
    }
}

static __device__ inline __attribute__((always_inline)) void plom_seir10_proj2obs_pp_f32(const float *X, const double *pvd, const double *lpd, const double *cvd, double *obs)
{
    (void)pvd; (void)lpd; (void)cvd; (void)X;
    obs[0] = (double)X[30];
    obs[1] = (double)X[31];
    obs[2] = (double)X[32];
    obs[3] = (double)X[33];
    obs[4] = (double)X[34];
    obs[5] = (double)X[35];
    obs[6] = (double)X[36];
    obs[7] = (double)X[37];
    obs[8] = (double)X[38];
    obs[9] = (double)X[39];
}

static __device__ inline __attribute__((always_inline)) double plom_seir10_loglik_pp(const double *obs, const double *y, const double *pvd, const double *lpd, const double *cvd, double t, double like_min, int *n_floor)
{
    double loglike = 0.0;
    int nf = 0;
    (void)t; (void)pvd; (void)lpd; (void)cvd;
    if (!plom_isnan(y[0])) {
        const double x = obs[0];
        const double gsl_mu = (lpd[50] * x);
        const double gsl_sd = plom_sqrt((((lpd[50] * (1.0 - lpd[50])) * x) + plom_powi(((lpd[50] * lpd[60]) * x), 2)));
        double like;
        if (y[0] > 0.0) like = plom_cdf_gaussian_P(y[0] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[0] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[0] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[1])) {
        const double x = obs[1];
        const double gsl_mu = (lpd[51] * x);
        const double gsl_sd = plom_sqrt((((lpd[51] * (1.0 - lpd[51])) * x) + plom_powi(((lpd[51] * lpd[61]) * x), 2)));
        double like;
        if (y[1] > 0.0) like = plom_cdf_gaussian_P(y[1] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[1] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[1] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[2])) {
        const double x = obs[2];
        const double gsl_mu = (lpd[52] * x);
        const double gsl_sd = plom_sqrt((((lpd[52] * (1.0 - lpd[52])) * x) + plom_powi(((lpd[52] * lpd[62]) * x), 2)));
        double like;
        if (y[2] > 0.0) like = plom_cdf_gaussian_P(y[2] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[2] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[2] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[3])) {
        const double x = obs[3];
        const double gsl_mu = (lpd[53] * x);
        const double gsl_sd = plom_sqrt((((lpd[53] * (1.0 - lpd[53])) * x) + plom_powi(((lpd[53] * lpd[63]) * x), 2)));
        double like;
        if (y[3] > 0.0) like = plom_cdf_gaussian_P(y[3] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[3] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[3] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[4])) {
        const double x = obs[4];
        const double gsl_mu = (lpd[54] * x);
        const double gsl_sd = plom_sqrt((((lpd[54] * (1.0 - lpd[54])) * x) + plom_powi(((lpd[54] * lpd[64]) * x), 2)));
        double like;
        if (y[4] > 0.0) like = plom_cdf_gaussian_P(y[4] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[4] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[4] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[5])) {
        const double x = obs[5];
        const double gsl_mu = (lpd[55] * x);
        const double gsl_sd = plom_sqrt((((lpd[55] * (1.0 - lpd[55])) * x) + plom_powi(((lpd[55] * lpd[65]) * x), 2)));
        double like;
        if (y[5] > 0.0) like = plom_cdf_gaussian_P(y[5] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[5] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[5] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[6])) {
        const double x = obs[6];
        const double gsl_mu = (lpd[56] * x);
        const double gsl_sd = plom_sqrt((((lpd[56] * (1.0 - lpd[56])) * x) + plom_powi(((lpd[56] * lpd[66]) * x), 2)));
        double like;
        if (y[6] > 0.0) like = plom_cdf_gaussian_P(y[6] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[6] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[6] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[7])) {
        const double x = obs[7];
        const double gsl_mu = (lpd[57] * x);
        const double gsl_sd = plom_sqrt((((lpd[57] * (1.0 - lpd[57])) * x) + plom_powi(((lpd[57] * lpd[67]) * x), 2)));
        double like;
        if (y[7] > 0.0) like = plom_cdf_gaussian_P(y[7] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[7] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[7] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[8])) {
        const double x = obs[8];
        const double gsl_mu = (lpd[58] * x);
        const double gsl_sd = plom_sqrt((((lpd[58] * (1.0 - lpd[58])) * x) + plom_powi(((lpd[58] * lpd[68]) * x), 2)));
        double like;
        if (y[8] > 0.0) like = plom_cdf_gaussian_P(y[8] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[8] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[8] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    if (!plom_isnan(y[9])) {
        const double x = obs[9];
        const double gsl_mu = (lpd[59] * x);
        const double gsl_sd = plom_sqrt((((lpd[59] * (1.0 - lpd[59])) * x) + plom_powi(((lpd[59] * lpd[69]) * x), 2)));
        double like;
        if (y[9] > 0.0) like = plom_cdf_gaussian_P(y[9] + 0.5 - gsl_mu, gsl_sd) - plom_cdf_gaussian_P(y[9] - 0.5 - gsl_mu, gsl_sd);
        else like = plom_cdf_gaussian_P(y[9] + 0.5 - gsl_mu, gsl_sd);
        loglike += plom_log(plom_sanitize_likelihood(like, like_min, &nf));
    }
    *n_floor = nf;
    return loglike;
}

struct plom_model_seir10 {
    static constexpr int S = 40, N_TS = 10, INC_OFFSET = 30, N_INC = 10, PV = 180, N_CAC = 10, N_HAT = 50, NPP = 70, FIX_OFF = 70, N_FIX = 10, N_DRIFT = 0;
    static constexpr int NTF = 0;  /* time-only functions of a substep */
    template <int V> static __device__ inline __attribute__((always_inline)) void interval(double *X, const double *pv, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
    {
        if constexpr (V == 0) plom_seir10_interval_full_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 1) plom_seir10_interval_no_dem_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 2) plom_seir10_interval_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 3) plom_seir10_interval_no_dem_sto_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 4) plom_seir10_interval_psr_full_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 5) plom_seir10_interval_psr_no_env_sto_f64(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
    }
    template <int V> static __device__ inline __attribute__((always_inline)) void interval_pp(double *X, const double *pv, const double *lp, plom_nstream64 *ns, double t0, double t1, double dt, double sqdt, const double *cv, int cs, const plom_drift_view *dm, const double *tft)
    {
        if constexpr (V == 0) plom_seir10_interval_full_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 1) plom_seir10_interval_no_dem_sto_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 2) plom_seir10_interval_no_env_sto_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 3) plom_seir10_interval_no_dem_sto_no_env_sto_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 4) plom_seir10_interval_psr_full_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 5) plom_seir10_interval_psr_no_env_sto_pp_f64(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
    }
    static __device__ inline __attribute__((always_inline)) void time_funcs(double tr, double *out) { plom_seir10_time_funcs_f64(tr, out); }
    static __device__ inline __attribute__((always_inline)) void proj2obs(const double *X, const double *pvd, const double *cvd, double *obs) { plom_seir10_proj2obs_f64(X, pvd, cvd, obs); }
    static __device__ inline __attribute__((always_inline)) void proj2obs_pp(const double *X, const double *pvd, const double *lpd, const double *cvd, double *obs) { plom_seir10_proj2obs_pp_f64(X, pvd, lpd, cvd, obs); }
    template <int V> static __device__ inline __attribute__((always_inline)) void interval(float *X, const float *pv, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
    {
        if constexpr (V == 0) plom_seir10_interval_full_f32(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 1) plom_seir10_interval_no_dem_sto_f32(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 2) plom_seir10_interval_no_env_sto_f32(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 3) plom_seir10_interval_no_dem_sto_no_env_sto_f32(X, pv, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
    }
    template <int V> static __device__ inline __attribute__((always_inline)) void interval_pp(float *X, const float *pv, const float *lp, plom_nstream32 *ns, double t0, double t1, double dt, float sqdt, const double *cv, int cs, const plom_drift_view *dm, const float *tft)
    {
        if constexpr (V == 0) plom_seir10_interval_full_pp_f32(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 1) plom_seir10_interval_no_dem_sto_pp_f32(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 2) plom_seir10_interval_no_env_sto_pp_f32(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
        else if constexpr (V == 3) plom_seir10_interval_no_dem_sto_no_env_sto_pp_f32(X, pv, lp, ns, t0, t1, dt, sqdt, cv, cs, dm, tft);
    }
    static __device__ inline __attribute__((always_inline)) void time_funcs(float tr, float *out) { plom_seir10_time_funcs_f32(tr, out); }
    static __device__ inline __attribute__((always_inline)) void proj2obs(const float *X, const double *pvd, const double *cvd, double *obs) { plom_seir10_proj2obs_f32(X, pvd, cvd, obs); }
    static __device__ inline __attribute__((always_inline)) void proj2obs_pp(const float *X, const double *pvd, const double *lpd, const double *cvd, double *obs) { plom_seir10_proj2obs_pp_f32(X, pvd, lpd, cvd, obs); }
    static __device__ inline __attribute__((always_inline)) int interval_ode(double *X, const double *pv, const double *cv, const plom_spline_set *sp, double t0, double t1, double dt, double ea, double er) { return plom_seir10_interval_ode_f64(X, pv, cv, sp, t0, t1, dt, ea, er); }
    static __device__ inline __attribute__((always_inline)) int interval_ode_pp(double *X, const double *pv, const double *lp, const double *cv, const plom_spline_set *sp, double t0, double t1, double dt, double ea, double er) { return plom_seir10_interval_ode_pp_f64(X, pv, lp, cv, sp, t0, t1, dt, ea, er); }
    static __device__ inline __attribute__((always_inline)) double loglik(const double *obs, const double *y, const double *pvd, const double *cvd, double t, double like_min, int *n_floor) { return plom_seir10_loglik(obs, y, pvd, cvd, t, like_min, n_floor); }
    static __device__ inline __attribute__((always_inline)) double loglik_pp(const double *obs, const double *y, const double *pvd, const double *lpd, const double *cvd, double t, double like_min, int *n_floor) { return plom_seir10_loglik_pp(obs, y, pvd, lpd, cvd, t, like_min, n_floor); }
    static __device__ inline __attribute__((always_inline)) void hat_values(const double *X, const double *pvd, const double *cvd, const plom_drift_view *dm, double t, double *out) { plom_seir10_hat_values(X, pvd, cvd, dm, t, out); }
    static __device__ inline __attribute__((always_inline)) void obs_moments(const double *obs, const double *pvd, const double *cvd, double t, double *mean, double *var) { plom_seir10_obs_moments(obs, pvd, cvd, t, mean, var); }
};
#endif

#endif
