// Device-side building blocks of the oxDNA2 hot path for gfx950:
// model constants, small FP32 / FP64 vectors, quaternion algebra and the
// Philox4x32-10 counter generator.  The modulation functions live in
// oxh_terms.inc with the pair terms.
//
// Model constants are the published oxDNA2 parameters as the reference
// writes them (src/model.h:15-434, src/defs.h:14 for the float PI); the
// modulation functions follow src/Interactions/DNAInteraction.cpp:1249-1454
// and src/Interactions/DNA2Interaction.cpp:308-363 (analytic f4, the
// reference's DNA2_nomesh form), evaluated in FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "oxh_types.h"

namespace oxh {

constexpr float PIf = 3.141592653589793238462643f;

constexpr float POS_BACK = -0.4f;
constexpr float POS_MM_BACK1 = -0.3400f;
constexpr float POS_MM_BACK2 = 0.3408f;
constexpr float POS_STACK = 0.34f;
constexpr float POS_BASE = 0.4f;
constexpr float GAMMA = POS_STACK - POS_BACK;

constexpr float FENE_EPS = 2.0f;
constexpr float FENE_DELTA = 0.25f;
constexpr float FENE_DELTA2 = 0.0625f;

constexpr float EXCL_EPS = 2.0f;
constexpr float EXCL_S1 = 0.70f, EXCL_S2 = 0.33f, EXCL_S3 = 0.515f, EXCL_S4 = 0.515f;
constexpr float EXCL_R1 = 0.675f, EXCL_R2 = 0.32f, EXCL_R3 = 0.50f, EXCL_R4 = 0.50f;
constexpr float EXCL_B1 = 892.016223343f, EXCL_B2 = 4119.70450017f, EXCL_B3 = 1707.30627298f, EXCL_B4 = 1707.30627298f;
constexpr float EXCL_RC1 = 0.711879214356f, EXCL_RC2 = 0.335388426126f, EXCL_RC3 = 0.52329943261f, EXCL_RC4 = 0.52329943261f;

constexpr float HYDR_RCLOW = 0.276908f, HYDR_RCHIGH = 0.783775f;
constexpr float CRST_RCLOW = 0.45f, CRST_RCHIGH = 0.7f;
constexpr float CXST_RCLOW = 0.177778f, CXST_RCHIGH = 0.6222222f;

// f1 radial: [0] HB, [1] stacking (src/model.h:92-102,145-157)
constexpr float F1_A[2] = {8.f, 6.f};
constexpr float F1_R0[2] = {0.4f, 0.4f};
constexpr float F1_BLOW[2] = {-126.243f, -68.1857f};
constexpr float F1_BHIGH[2] = {-7.87708f, -3.12992f};
constexpr float F1_RLOW[2] = {0.34f, 0.32f};
constexpr float F1_RHIGH[2] = {0.7f, 0.75f};
constexpr float F1_RCLOW[2] = {0.276908f, 0.23239f};
constexpr float F1_RCHIGH[2] = {0.783775f, 0.956f};

// f2 radial: [0] cross stacking, [1] coaxial stacking, oxDNA2 K
// (src/model.h:193-202,349-360; src/Interactions/DNA2Interaction.cpp:10)
constexpr float F2_K[2] = {47.5f, 58.5f};
constexpr float F2_RC[2] = {0.675f, 0.6f};
constexpr float F2_R0[2] = {0.575f, 0.400f};
constexpr float F2_BLOW[2] = {-0.888889f, -2.13158f};
constexpr float F2_BHIGH[2] = {-0.888889f, -2.13158f};
constexpr float F2_RLOW[2] = {0.495f, 0.22f};
constexpr float F2_RHIGH[2] = {0.655f, 0.58f};
constexpr float F2_RCLOW[2] = {0.45f, 0.177778f};
constexpr float F2_RCHIGH[2] = {0.7f, 0.6222222f};

// f4 slots, numbered as the reference numbers them (src/model.h:105-110,...)
enum {
	S_STCK_T4 = 0, S_STCK_T5 = 1, S_HYDR_T1 = 2, S_HYDR_T2 = 3, S_HYDR_T4 = 4, S_HYDR_T7 = 5,
	S_CRST_T1 = 6, S_CRST_T2 = 7, S_CRST_T4 = 8, S_CRST_T7 = 9, S_CXST_T1 = 10, S_CXST_T4 = 11, S_CXST_T5 = 12
};
constexpr float F4_A[13] = {1.3f, 0.9f, 1.5f, 1.5f, 0.46f, 4.f, 2.25f, 1.70f, 1.50f, 1.70f, 2.f, 1.3f, 0.9f};
constexpr float F4_B[13] = {6.4381f, 3.89361f, 4.16038f, 4.16038f, 0.133855f, 17.0526f, 7.00545f, 6.2469f, 2.59556f, 6.2469f,
		10.9032f, 6.4381f, 3.89361f};
constexpr float F4_T0[13] = {0.f, 0.f, 0.f, 0.f, PIf, PIf * 0.5f, PIf - 2.35f, 1.f, 0.f, 0.875f, PIf - 0.25f, 0.f, 0.f};
constexpr float F4_TS[13] = {0.8f, 0.95f, 0.7f, 0.7f, 0.7f, 0.45f, 0.58f, 0.68f, 0.65f, 0.68f, 0.65f, 0.8f, 0.95f};
constexpr float F4_TC[13] = {0.961538f, 1.16959f, 0.952381f, 0.952381f, 3.10559f, 0.555556f, 0.766284f, 0.865052f, 1.02564f,
		0.865052f, 0.769231f, 0.961538f, 1.16959f};
constexpr float CXST_THETA1_SA = 20.f;
constexpr float CXST_THETA1_SB = (PIf - 0.1f * (PIf - (PIf - 0.25f)));

// f5 (phi) modulation (src/model.h:180-189)
constexpr float F5_A = 2.0f, F5_B = 10.9032f, F5_XC = -0.769231f, F5_XS = -0.65f;

// centre distance beyond which only Debye-Hueckel can act: the largest
// site-pair range (back-back excluded volume, 0.7119) plus twice the
// backbone-site arm |(-0.34, 0.3408)| = 0.4814
constexpr float R_CLOSE = 1.676f;
// workgroup size of the packed form of the close kernel
constexpr int PACK_BS = 256;

// ------------------------------------------------------------ small vectors
struct f3 {
	float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 &operator+=(f3 &a, f3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
__device__ __forceinline__ f3 &operator-=(f3 &a, f3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ __forceinline__ float SQR(float x) { return x * x; }

// FP64 small vectors, same operations
struct dv3 {
	double x, y, z;
};
__device__ __forceinline__ dv3 mk(double x, double y, double z) { return {x, y, z}; }
__device__ __forceinline__ dv3 operator+(dv3 a, dv3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ dv3 operator-(dv3 a, dv3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ dv3 operator-(dv3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ dv3 operator*(dv3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ dv3 operator*(double s, dv3 a) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ dv3 &operator+=(dv3 &a, dv3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
__device__ __forceinline__ dv3 &operator-=(dv3 &a, dv3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
__device__ __forceinline__ double dot(dv3 a, dv3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ dv3 cross(dv3 a, dv3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ __forceinline__ double SQR(double x) { return x * x; }

// body axes from a unit quaternion (x, y, z, w): a1, a2, a3 are the images
// of the x, y, z axes (the columns of the rotation matrix)
__device__ __forceinline__ void quat_axes(float4 q, f3 &a1, f3 &a2, f3 &a3) {
	float sqx = q.x * q.x, sqy = q.y * q.y, sqz = q.z * q.z, sqw = q.w * q.w;
	float xy = q.x * q.y, xz = q.x * q.z, xw = q.x * q.w, yz = q.y * q.z, yw = q.y * q.w, zw = q.z * q.w;
	a1 = {sqx - sqy - sqz + sqw, 2.f * (xy + zw), 2.f * (xz - yw)};
	a2 = {2.f * (xy - zw), -sqx + sqy - sqz + sqw, 2.f * (yz + xw)};
	a3 = {2.f * (xz + yw), 2.f * (yz - xw), -sqx - sqy + sqz + sqw};
}

// the same axes in FP64 from an FP32 quaternion
__device__ __forceinline__ void quat_axes(float4 qf, dv3 &a1, dv3 &a2, dv3 &a3) {
	const double qx = qf.x, qy = qf.y, qz = qf.z, qw = qf.w;
	double sqx = qx * qx, sqy = qy * qy, sqz = qz * qz, sqw = qw * qw;
	double xy = qx * qy, xz = qx * qz, xw = qx * qw, yz = qy * qz, yw = qy * qw, zw = qz * qw;
	a1 = {sqx - sqy - sqz + sqw, 2. * (xy + zw), 2. * (xz - yw)};
	a2 = {2. * (xy - zw), -sqx + sqy - sqz + sqw, 2. * (yz + xw)};
	a3 = {2. * (xz + yw), 2. * (yz - xw), -sqx - sqy + sqz + sqw};
}

// ------------------------------------------------------------ Philox4x32-10
// J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11.
__host__ __device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k) {
#pragma unroll
	for(int r = 0; r < 10; r++) {
		uint64_t p0 = (uint64_t) 0xD2511F53u * c.x;
		uint64_t p1 = (uint64_t) 0xCD9E8D57u * c.z;
		uint32_t hi0 = (uint32_t) (p0 >> 32), lo0 = (uint32_t) p0;
		uint32_t hi1 = (uint32_t) (p1 >> 32), lo1 = (uint32_t) p1;
		c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
		k.x += 0x9E3779B9u;
		k.y += 0xBB67AE85u;
	}
	return c;
}

// uniform in (0, 1) from a 32-bit word
__host__ __device__ __forceinline__ double u01(uint32_t x) { return ((double) x + 0.5) * 2.3283064365386963e-10; }

// two standard normals (Box-Muller, FP64) from two 32-bit words
__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, double &n0, double &n1) {
	double r = sqrt(-2.0 * log(u01(a)));
	double s, c;
	sincospi(2.0 * u01(b), &s, &c);
	n0 = r * c;
	n1 = r * s;
}

// Workgroups of a launch go to the chip's 8 XCDs in turn (workgroup b runs
// on XCD b % 8), and each XCD has its own L2.  A launch whose workgroups
// gather their neighbours' data takes its grid as xcd_grid(n) and its
// workgroup index as xcd_block(n): XCD x then runs the contiguous logical
// blocks [x ceil(n / 8), (x + 1) ceil(n / 8)), so particles that are close
// in memory order (strand by strand, or Hilbert-sorted) and in space are
// gathered through one L2.  -1: a padding workgroup of the rounded-up grid.
constexpr unsigned N_XCD = 8;
__host__ __device__ __forceinline__ unsigned xcd_grid(unsigned n) { return (n + N_XCD - 1) / N_XCD * N_XCD; }
__device__ __forceinline__ int xcd_block(unsigned n) {
	const unsigned per = (n + N_XCD - 1) / N_XCD;
	const unsigned lb = (blockIdx.x % N_XCD) * per + blockIdx.x / N_XCD;
	return lb < n ? (int) lb : -1;
}

// particle i is padding of a slot: beyond its system's current size
__device__ __forceinline__ bool padding(int i, const oxh_sys &S) { return i - S.first >= S.n; }

// particle i takes no part in step k of the chunk: padding, or its system
// has used up its steps (limit[s], slot pools)
__device__ __forceinline__ bool idle(int i, const oxh_sys &S, int s, int k, const int *__restrict__ limit) {
	return padding(i, S) || (limit && k >= limit[s]);
}

}  // namespace oxh
