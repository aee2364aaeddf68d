// Pair physics of the oxDNA2 interaction on gfx950, shared by the grid
// force kernels (oxh_forces.hip) and the workgroup-resident engine
// (oxh_persist.hip): the FP64 site geometry, and the pair terms of
// oxh_terms.inc in FP32 (namespace sp, the default) and FP64 (namespace
// dp, the all-FP64 mode).
#pragma once

#include "oxh_device.h"

namespace oxh {

// Work lists of the packed form of k_forces (one per workgroup of
// PACK_BS threads, in LDS): the pairs of one round whose base-base
// separation reaches HB or CRST (h) and those whose stack-stack separation
// reaches CXST (c), one item per slot, fields as separate arrays so that
// consecutive slots are conflict-free.  Fields of h: p's a1, a3, q's a1, a3,
// p's base offset, base-base separation (18 floats), p's and q's base type
// and the system (ints).  Fields of c: p's a1, a3, q's a1, a3, p's stack
// offset, stack-stack separation.  Evaluating an item overwrites its first
// fields with the results: force on q, lab torque on p, energies.
constexpr int PACK_H = 21, PACK_C = 18;
struct PackShared {
	float h[PACK_H][PACK_BS];
	float c[PACK_C][PACK_BS];
	int nh[PACK_BS / 64], nc[PACK_BS / 64], rounds[PACK_BS / 64];
};

// a nucleotide's FP64 a1 axis and backbone offset (the stacking and base
// offsets are POS_STACK a1 and POS_BASE a1)
struct Sites64 {
	dv3 a1, back;
};

__device__ __forceinline__ void sites_d(double4 q, Sites64 &st) {
	const double sqx = q.x * q.x, sqy = q.y * q.y, sqz = q.z * q.z, sqw = q.w * q.w;
	const double xy = q.x * q.y, xz = q.x * q.z, xw = q.x * q.w, yz = q.y * q.z, yw = q.y * q.w, zw = q.z * q.w;
	const dv3 a1 = {sqx - sqy - sqz + sqw, 2. * (xy + zw), 2. * (xz - yw)};
	const dv3 a2 = {2. * (xy - zw), -sqx + sqy - sqz + sqw, 2. * (yz + xw)};
	const double b1 = (double) POS_MM_BACK1, b2 = (double) POS_MM_BACK2;
	st.a1 = a1;
	st.back = {b1 * a1.x + b2 * a2.x, b1 * a1.y + b2 * a2.y, b1 * a1.z + b2 * a2.z};
}

namespace sp {
using real = float;
using vec = f3;
using out4 = float4;
constexpr bool FP64_TERMS = false;
__device__ __forceinline__ real pick(float f, double) { return f; }
__device__ __forceinline__ out4 make_out4(real x, real y, real z, real w) { return make_float4(x, y, z, w); }
// |r| and 1 / |r| from r^2 by the hardware reciprocal square root (~1 ulp),
// and e^x by the hardware base-2 exponential (relative error ~|x| ulp):
// Debye-Hueckel's far pairs outnumber the close ones tenfold in dense
// systems, and these replace a correctly rounded sqrt, divide and exp
__device__ __forceinline__ void norm_inv(real r2, real &r, real &inv) {
	inv = rsqrtf(r2);
	r = r2 * inv;
}
__device__ __forceinline__ real exp_dh(real x) { return __expf(x); }
#include "oxh_terms.inc"
}  // namespace sp

namespace dp {
using real = double;
using vec = dv3;
using out4 = double4;
constexpr bool FP64_TERMS = true;
__device__ __forceinline__ real pick(float, double d) { return d; }
__device__ __forceinline__ out4 make_out4(real x, real y, real z, real w) { return make_double4(x, y, z, w); }
__device__ __forceinline__ void norm_inv(real r2, real &r, real &inv) {
	r = sqrt(r2);
	inv = real(1.) / r;
}
__device__ __forceinline__ real exp_dh(real x) { return exp(x); }
#include "oxh_terms.inc"
}  // namespace dp

// the FP32 terms are the default
using sp::add_virial;
using sp::debye_huckel;
using sp::Frame;
using sp::group_sum;
using sp::make_frame;
using sp::min_image;
using sp::min_image_f;
using sp::particle_forces;
using sp::edge_item;
using sp::make_frame_d;

}  // namespace oxh
