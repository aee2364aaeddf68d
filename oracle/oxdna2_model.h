/*
 * oxDNA2 model constants, shared by the FP64 oracle (oracle/oxdna2_oracle.cpp)
 * and, as float values, by the HIP kernels (oxdna_amd/csrc/oxh_model.h keeps
 * its own copy so the product never includes anything from oracle/).
 *
 * The values are the published oxDNA/oxDNA2 model parameters as the
 * reference writes them in src/model.h:15-434.  The reference spells every
 * constant as a *float* literal and `PI` as a float literal too
 * (src/defs.h:14), and stores them into FP64 `number` members; an FP64
 * restatement that means to agree with the reference to 1e-12 has to start
 * from the same float-rounded values, so they are written here as float
 * literals and converted to double exactly once.  Composite angles such as
 * (PI - 2.35f) are evaluated in float arithmetic, as C++ evaluates them in
 * the reference.
 */
#ifndef OXDNA2_MODEL_H
#define OXDNA2_MODEL_H

namespace oxmodel {

constexpr float PIf = 3.141592653589793238462643f;

// interaction-site offsets along the body axes (src/model.h:15-20)
constexpr float POS_BACK = -0.4f;
constexpr float POS_MM_BACK1 = -0.3400f;
constexpr float POS_MM_BACK2 = 0.3408f;
constexpr float POS_STACK = 0.34f;
constexpr float POS_BASE = 0.4f;

// FENE (src/model.h:47-52)
constexpr float FENE_EPS = 2.0f;
constexpr float FENE_R0_OXDNA2 = 0.7564f;
constexpr float FENE_DELTA = 0.25f;
constexpr float FENE_DELTA2 = 0.0625f;

// excluded volume (src/model.h:57-85); index 1..7 as in the reference
constexpr float EXCL_EPS = 2.0f;
constexpr float EXCL_S1 = 0.70f, EXCL_S2 = 0.33f, EXCL_S3 = 0.515f, EXCL_S4 = 0.515f;
constexpr float EXCL_R1 = 0.675f, EXCL_R2 = 0.32f, EXCL_R3 = 0.50f, EXCL_R4 = 0.50f;
constexpr float EXCL_B1 = 892.016223343f, EXCL_B2 = 4119.70450017f, EXCL_B3 = 1707.30627298f, EXCL_B4 = 1707.30627298f;
constexpr float EXCL_RC1 = 0.711879214356f, EXCL_RC2 = 0.335388426126f, EXCL_RC3 = 0.52329943261f, EXCL_RC4 = 0.52329943261f;

// hydrogen bonding, radial (src/model.h:92-102)
constexpr float HYDR_EPS_OXDNA2 = 1.0678f;
constexpr float HYDR_A = 8.f, HYDR_RC = 0.75f, HYDR_R0 = 0.4f;
constexpr float HYDR_BLOW = -126.243f, HYDR_BHIGH = -7.87708f;
constexpr float HYDR_RLOW = 0.34f, HYDR_RHIGH = 0.7f, HYDR_RCLOW = 0.276908f, HYDR_RCHIGH = 0.783775f;

// stacking, radial (src/model.h:145-157)
constexpr float STCK_BASE_EPS_OXDNA2 = 1.3523f, STCK_FACT_EPS_OXDNA2 = 2.6717f;
constexpr float STCK_A = 6.f, STCK_RC = 0.9f, STCK_R0 = 0.4f;
constexpr float STCK_BLOW = -68.1857f, STCK_BHIGH = -3.12992f;
constexpr float STCK_RLOW = 0.32f, STCK_RHIGH = 0.75f, STCK_RCLOW = 0.23239f, STCK_RCHIGH = 0.956f;

// cross stacking, radial (src/model.h:193-202)
constexpr float CRST_R0 = 0.575f, CRST_RC = 0.675f, CRST_K = 47.5f;
constexpr float CRST_BLOW = -0.888889f, CRST_RLOW = 0.495f, CRST_RCLOW = 0.45f;
constexpr float CRST_BHIGH = -0.888889f, CRST_RHIGH = 0.655f, CRST_RCHIGH = 0.7f;

// coaxial stacking, radial (src/model.h:349-360)
constexpr float CXST_R0 = 0.400f, CXST_RC = 0.6f, CXST_K_OXDNA2 = 58.5f;
constexpr float CXST_BLOW = -2.13158f, CXST_RLOW = 0.22f, CXST_RCLOW = 0.177778f;
constexpr float CXST_BHIGH = -2.13158f, CXST_RHIGH = 0.58f, CXST_RCHIGH = 0.6222222f;

// f1 families: 0 = HB, 1 = stacking
enum { F1_HYDR = 0, F1_STCK = 1 };
// f2 families: 0 = cross stacking, 1 = coaxial stacking
enum { F2_CRST = 0, F2_CXST = 1 };

/* f4 angular modulation slots, numbered as the reference numbers them
 * (src/model.h:105-110,160-162,204-209,362-365). */
enum {
	F4_STCK_T4 = 0, F4_STCK_T5 = 1,                 // t5 and t6 share slot 1
	F4_HYDR_T1 = 2, F4_HYDR_T2 = 3, F4_HYDR_T4 = 4, F4_HYDR_T7 = 5,   // t3 -> 3, t8 -> 5
	F4_CRST_T1 = 6, F4_CRST_T2 = 7, F4_CRST_T4 = 8, F4_CRST_T7 = 9,   // t3 -> 7, t8 -> 9
	F4_CXST_T1 = 10, F4_CXST_T4 = 11, F4_CXST_T5 = 12,               // t6 -> 12
	F4_N = 13
};

// per-slot (A, B, T0, TS, TC); T0 entries that involve PI are float
// expressions, exactly as the reference's macros expand
// (src/model.h:112-140,163-177,210-240,366-387; DNA2 overrides slot 10's T0,
// src/Interactions/DNA2Interaction.cpp:12)
constexpr float F4_A[F4_N] = {1.3f, 0.9f, 1.5f, 1.5f, 0.46f, 4.f, 2.25f, 1.70f, 1.50f, 1.70f, 2.f, 1.3f, 0.9f};
constexpr float F4_B[F4_N] = {6.4381f, 3.89361f, 4.16038f, 4.16038f, 0.133855f, 17.0526f, 7.00545f, 6.2469f, 2.59556f, 6.2469f,
		10.9032f, 6.4381f, 3.89361f};
constexpr float F4_T0[F4_N] = {0.f, 0.f, 0.f, 0.f, PIf, PIf * 0.5f, PIf - 2.35f, 1.f, 0.f, 0.875f, PIf - 0.25f, 0.f, 0.f};
constexpr float F4_TS[F4_N] = {0.8f, 0.95f, 0.7f, 0.7f, 0.7f, 0.45f, 0.58f, 0.68f, 0.65f, 0.68f, 0.65f, 0.8f, 0.95f};
constexpr float F4_TC[F4_N] = {0.961538f, 1.16959f, 0.952381f, 0.952381f, 3.10559f, 0.555556f, 0.766284f, 0.865052f, 1.02564f,
		0.865052f, 0.769231f, 0.961538f, 1.16959f};

// oxDNA2 extra harmonic term on coaxial theta1 (src/model.h:371-372)
constexpr float CXST_THETA1_SA = 20.f;
constexpr float CXST_THETA1_SB = (PIf - 0.1f * (PIf - (PIf - 0.25f)));

// f5 (phi) modulation slots: 0 = stck phi1, 1 = stck phi2, 2/3 = cxst phi3/phi4
// (src/model.h:180-189,389-398; slot 3 takes PHI3's B as the reference does,
// src/Interactions/DNAInteraction.cpp:197)
constexpr float F5_A[4] = {2.0f, 2.0f, 2.0f, 2.0f};
constexpr float F5_B[4] = {10.9032f, 10.9032f, 10.9032f, 10.9032f};
constexpr float F5_XC[4] = {-0.769231f, -0.769231f, -0.769231f, -0.769231f};
constexpr float F5_XS[4] = {-0.65f, -0.65f, -0.65f, -0.65f};

// Debye-Hueckel defaults (src/Interactions/DNA2Interaction.cpp:68-76)
constexpr double DH_LAMBDA_DEFAULT = 0.3616455;
constexpr double DH_STRENGTH_DEFAULT = 0.0543;

// hydrogen-bond cut-off used by the base-pair counter (src/defs.h:62)
constexpr float HB_CUTOFF = -0.1f;

}  // namespace oxmodel

#endif
