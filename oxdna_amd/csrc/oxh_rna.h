// The oxRNA model (RNA and RNA2) on gfx950: its constants.
//
// The constants are the defaults of the reference's Model struct
// (src/Interactions/rna_model.h:311-654), which holds them as floats; each
// is written here as the reference assigns it (a double literal rounded to
// float, or a float literal), and the terms promote them to double as the
// reference's double build does.  The indices of the f1, f2, f4 and f5
// tables are the reference's (rna_model.h:16-52: RNA_HYDR_F1 = 0,
// RNA_STCK_F1 = 1, RNA_CRST_F2 = 0, RNA_CXST_F2 = 1, and the f4 / f5 type
// numbers), filled as RNAInteraction::init does
// (src/Interactions/RNAInteraction.cpp:110-294), including its reuse of
// one entry for two angles and F5_PHI_B[3] = RNA_CXST_PHI3_B.
//
// The hydrogen-bonding f4 factors are the analytic ones, as the reference's
// GPU kernel evaluates them (src/CUDA/Interactions/CUDA_RNA.cuh:743-761).
// Its CPU backend interpolates them instead on cubic meshes of 6 to 250
// points in cos(theta) (RNAInteraction.cpp:312-324, Mesh.cpp:30-56), which
// departs from the analytic form by up to a few 1e-3 per factor.
#pragma once

#include "oxh_device.h"

namespace oxh {
namespace rna {

#define RF(x) ((float) (x))

// interaction sites in the body frame (RNANucleotide.h:25-47)
constexpr float POS_BACK_A1 = RF(-0.4), POS_BACK_A2 = RF(0.0), POS_BACK_A3 = RF(0.2);
constexpr float POS_STACK = 0.34f, POS_BASE = 0.4f;
constexpr float POS_STACK3_A1 = RF(0.4), POS_STACK3_A2 = RF(0.1);
constexpr float POS_STACK5_A1 = RF(0.124906078525), POS_STACK5_A2 = RF(-0.00866274917473);
constexpr float P3_X = RF(-0.462510), P3_Y = RF(-0.528218), P3_Z = RF(0.712089);
constexpr float P5_X = RF(-0.104402), P5_Y = RF(-0.841783), P5_Z = RF(0.529624);

// FENE and excluded volume
constexpr float FENE_EPS = 2.0f, FENE_R0 = RF(0.761070781051), FENE_DELTA = 0.25f, FENE_DELTA2 = 0.0625f;
constexpr float EXCL_EPS = 2.0f;
constexpr float EXCL_S1 = 0.70f, EXCL_S2 = 0.33f, EXCL_S3 = 0.515f, EXCL_S4 = 0.515f;
constexpr float EXCL_R1 = 0.675f, EXCL_R2 = 0.32f, EXCL_R3 = 0.50f, EXCL_R4 = 0.50f;
constexpr float EXCL_B1 = 892.016223343f, EXCL_B2 = 4119.70450017f, EXCL_B3 = 1707.30627298f, EXCL_B4 = 1707.30627298f;
constexpr float EXCL_RC1 = 0.711879214356f, EXCL_RC2 = 0.335388426126f, EXCL_RC3 = 0.52329943261f, EXCL_RC4 = 0.52329943261f;

// f1: 0 hydrogen bonding, 1 stacking
constexpr float HYDR_EPS = 0.870439f;
constexpr float STCK_BASE_EPS = 1.40206f, STCK_FACT_EPS = 2.77f;
constexpr float F1_A[2] = {8.f, 6.f};
constexpr float F1_RC[2] = {0.75f, RF(0.93)};
constexpr float F1_R0[2] = {0.4f, RF(0.43)};
constexpr float F1_BLOW[2] = {-126.243f, -68.1857f};
constexpr float F1_BHIGH[2] = {-7.87708f, -3.12992f};
constexpr float F1_RLOW[2] = {0.34f, RF(0.35)};
constexpr float F1_RHIGH[2] = {0.7f, RF(0.78)};
constexpr float F1_RCLOW[2] = {0.276908f, RF(0.26239)};
constexpr float F1_RCHIGH[2] = {0.783775f, RF(0.986)};
constexpr float HYDR_RCLOW = F1_RCLOW[0], HYDR_RCHIGH = F1_RCHIGH[0];

// f2: 0 cross stacking, 1 coaxial stacking
constexpr float F2_K[2] = {RF(59.9626), RF(80.)};
constexpr float F2_RC[2] = {RF(0.6), RF(0.6)};
constexpr float F2_R0[2] = {RF(0.5), RF(0.5)};
constexpr float F2_BLOW[2] = {-0.888889f, RF(-0.888889)};
constexpr float F2_BHIGH[2] = {-0.888889f, RF(-0.888889)};
constexpr float F2_RLOW[2] = {RF(0.42), RF(0.42)};
constexpr float F2_RHIGH[2] = {RF(0.58), RF(0.58)};
constexpr float F2_RCLOW[2] = {RF(0.375), RF(0.375)};
constexpr float F2_RCHIGH[2] = {RF(0.625), RF(0.625)};

// f4 types (rna_model.h:19-49)
constexpr int STCK_T5 = 1, STCK_T6 = 13, STCK_TB1 = 14, STCK_TB2 = 15;
constexpr int HYDR_T1 = 2, HYDR_T2 = 3, HYDR_T3 = 3, HYDR_T4 = 4, HYDR_T7 = 5, HYDR_T8 = 5;
constexpr int CRST_T1 = 6, CRST_T2 = 7, CRST_T3 = 7, CRST_T4 = 8, CRST_T7 = 9, CRST_T8 = 9;
constexpr int CXST_T1 = 10, CXST_T4 = 11, CXST_T5 = 12, CXST_T6 = 12;
//                                  0 stck t4  1 stck t5  2 hydr t1  3 hydr t2  4 hydr t4  5 hydr t7  6 crst t1  7 crst t2
//                                  8 crst t4  9 crst t7  10 cxst t1 11 cxst t4 12 cxst t5 13 stck t6  14 stck tB1 15 stck tB2
constexpr float F4_A[16] = {RF(0), 0.9f, 1.5f, 1.5f, 0.46f, 4.f, 2.25f, 1.70f,
                            1.50f, 1.70f, 2.f, 1.3f, 0.9f, 0.9f, RF(1.3), RF(1.3)};
constexpr float F4_B[16] = {RF(0), 3.89361f, 4.16038f, 4.16038f, 0.133855f, 17.0526f, 7.00545f, 6.2469f,
                            2.59556f, 6.2469f, 10.9032f, 6.4381f, 3.89361f, 3.89361f, RF(6.4381), RF(6.4381)};
constexpr float F4_T0[16] = {0.f, 0.f, 0.f, 0.f, PIf, PIf * 0.5f, RF(0.505), RF(1.266),
                             RF(0), RF(0.309), RF(2.592), RF(0.151), RF(0.685), 0.f, 0.f, 0.f};
constexpr float F4_TS[16] = {0.8f, 0.95f, 0.7f, 0.7f, 0.7f, 0.45f, 0.58f, 0.68f,
                             0.65f, 0.68f, 0.65f, 0.8f, 0.95f, 0.95f, RF(0.8), RF(0.8)};
constexpr float F4_TC[16] = {0.961538f, 1.16959f, 0.952381f, 0.952381f, 3.10559f, 0.555556f, 0.766284f, 0.865052f,
                             1.02564f, 0.865052f, 0.769231f, 0.961538f, 1.16959f, 1.16959f, RF(0.961538), RF(0.961538)};

// f5 types: 0 stck phi1, 1 stck phi2, 2 cxst phi3, 3 cxst phi4
constexpr int STCK_PHI1 = 0, STCK_PHI2 = 1, CXST_PHI3 = 2, CXST_PHI4 = 3;
constexpr float F5_A[4] = {2.0f, 2.0f, 2.0f, 2.0f};
constexpr float F5_B[4] = {10.9032f, 10.9032f, 10.9032f, 10.9032f};
constexpr float F5_XC[4] = {-0.769231f, -0.769231f, -0.769231f, -0.769231f};
constexpr float F5_XS[4] = {-0.65f, -0.65f, -0.65f, -0.65f};

#undef RF

}  // namespace rna
}  // namespace oxh
