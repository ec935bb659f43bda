/* Generated by gen_phi_table.py: Chebyshev coefficients for plom_Phi. Do not edit. */
#ifndef PLOM_PHI_COEFFS_H
#define PLOM_PHI_COEFFS_H
/* central: variable in [0.0, 1.0], degree 14, last coefficient 5.68e-26 */
#define PLOM_PHI_CENTRAL_LO 0.0
#define PLOM_PHI_CENTRAL_HI 1.0
#define PLOM_PHI_CENTRAL_DEG 14
#define PLOM_PHI_CENTRAL_COEFFS { 0x1.79f42ccb14b3ap-2, -0x1.d75627af43af0p-6, 0x1.12886985bb256p-10, -0x1.01220e24e861ap-15, 0x1.8bbc91d1c9165p-21, -0x1.011d4ea74b888p-26, 0x1.2080afd361863p-32, -0x1.1c9108887c605p-38, 0x1.f4868f67e6cfbp-45, -0x1.8d04d74ce627cp-51, 0x1.1ebbd2d2ea8fep-57, -0x1.7c18a8ad9cb82p-64, 0x1.d184536c49a67p-71, -0x1.08e4372b0199cp-77, 0x1.1981601262f95p-84 }
/* mills_1_2: variable in [0.5, 1.0], degree 22, last coefficient -2.91e-21 */
#define PLOM_PHI_MILLS_1_2_LO 0.5
#define PLOM_PHI_MILLS_1_2_HI 1.0
#define PLOM_PHI_MILLS_1_2_DEG 22
#define PLOM_PHI_MILLS_1_2_COEFFS { 0x1.7da6cda6e512fp-1, -0x1.7fa29ad8a6f86p-4, 0x1.f8be04ea4335ep-9, 0x1.024f5f02836e6p-13, -0x1.c6917e0f05fdep-15, 0x1.13e7c9104b703p-17, -0x1.ecc5c00e34b74p-21, 0x1.53deeb0fdcbbep-24, -0x1.229c5c0b9e58ep-28, -0x1.8add1bc9c91abp-33, 0x1.aa8402fabe834p-34, -0x1.45f4dff298f7cp-36, 0x1.78d0be9c785a8p-39, -0x1.6c0b897016a76p-42, 0x1.25768b6601975p-45, -0x1.6201fb42f4c58p-49, 0x1.f09d879ee6abfp-55, 0x1.f5b6a6aac16adp-56, -0x1.109b15b791326p-57, 0x1.87b9e2e24e2efp-60, -0x1.d2ba84d2b9cfdp-63, 0x1.e5bb8d6a07f29p-66, -0x1.b77bbae8a233dp-69 }
/* mills_2_4: variable in [0.25, 0.5], degree 20, last coefficient 9.77e-22 */
#define PLOM_PHI_MILLS_2_4_LO 0.25
#define PLOM_PHI_MILLS_2_4_HI 0.5
#define PLOM_PHI_MILLS_2_4_DEG 20
#define PLOM_PHI_MILLS_2_4_COEFFS { 0x1.ca80d9bece166p-1, -0x1.aba0a085c36b1p-5, -0x1.ad2b0f945da50p-11, 0x1.151825272944fp-12, -0x1.6d6e6fd1833adp-16, 0x1.c00f82cd9eb94p-21, 0x1.91ae37d207dd7p-25, -0x1.a988505e8a0c7p-27, 0x1.6f4f814fb12f7p-30, -0x1.7e946607f77a4p-34, 0x1.0bd384dab3bc8p-40, 0x1.812b348626f7bp-41, -0x1.0e832102da3a5p-43, 0x1.d036e209eea56p-47, -0x1.0681394851cdbp-50, 0x1.38005436b16e7p-56, 0x1.0be641f27a56bp-57, -0x1.bc820b01a8298p-60, 0x1.c285fc0571998p-63, -0x1.498629a7bf9d9p-66, 0x1.2766cadcb76bcp-70 }
/* mills_4_8: variable in [0.125, 0.25], degree 16, last coefficient -4.15e-20 */
#define PLOM_PHI_MILLS_4_8_LO 0.125
#define PLOM_PHI_MILLS_4_8_HI 0.25
#define PLOM_PHI_MILLS_4_8_DEG 16
#define PLOM_PHI_MILLS_4_8_COEFFS { 0x1.ef113fb2671f9p-1, -0x1.3c06aa9fa17b0p-6, -0x1.1f1f65aad26e6p-10, 0x1.13cae6071ae7ep-14, -0x1.36899acb1fa07p-21, -0x1.319e37bd337f0p-23, 0x1.605b7156c9103p-27, -0x1.7f2b8b99124adp-33, -0x1.eac96d6632a5bp-36, 0x1.9c173dc800b78p-39, -0x1.fdcbf2e972c43p-44, -0x1.4fb7b1f1263d5p-48, 0x1.28af690ec0c45p-50, -0x1.5260e8dc7560cp-54, 0x1.8f80fed390996p-60, 0x1.69b3a537648c3p-62, -0x1.87ec29c04db97p-65 }
/* mills_8_40: variable in [0.025, 0.125], degree 18, last coefficient -1.78e-22 */
#define PLOM_PHI_MILLS_8_40_LO 0.025
#define PLOM_PHI_MILLS_8_40_HI 0.125
#define PLOM_PHI_MILLS_8_40_DEG 18
#define PLOM_PHI_MILLS_8_40_COEFFS { 0x1.fc97144114492p-1, -0x1.d6d3b3b05d5dep-8, -0x1.2718d7c5ac92ep-10, 0x1.9476d7ddb3b5ap-16, 0x1.93554c26a8f1dp-20, -0x1.82ff95f7ce6aap-24, -0x1.d26b31960af10p-30, 0x1.847da1909d66ep-32, -0x1.c3e42aecff9d5p-38, -0x1.619262a354f91p-40, 0x1.855e0f4d45a3ep-44, 0x1.56f1838fb119ap-49, -0x1.61c2a804c257bp-51, 0x1.69fbb5e32292ep-56, 0x1.a3df8f2cc4503p-59, -0x1.73099787dd181p-62, -0x1.ead70591933ebp-73, 0x1.72294dcf5d30bp-69, -0x1.ad7f4c13625efp-73 }
#endif
