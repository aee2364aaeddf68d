// Per-pair quantities of FFS order parameters on gfx950, in FP64.
//
// The reference evaluates its order parameters (src/Utilities/
// OrderParameters.cpp) on the host: a "bond" parameter counts its listed
// pairs whose hydrogen-bonding energy is below HB_CUTOFF (-0.1,
// src/Observables/OrderParameterValues.cpp:55-66), a "mindistance"
// parameter takes the smallest base-base (or centre-centre) distance of
// its listed pairs, the plain difference for pairs of one strand and the
// nearest image otherwise (OrderParameters.cpp:73-98).  Its CUDA FFS
// backend precomputes the same per-pair quantities on the GPU
// (hb_op_precalc, near_hb_op_precalc, dist_op_precalc,
// src/CUDA/Interactions/CUDA_DNA.cuh:909-1069), one thread per listed pair;
// so does this kernel, from the batch's FP64 state.  The host folds the
// per-pair values into parameters.
#include "../../include/oxh.h"
#include "oxh_device.h"
#include "oxh_internal.h"
#include "oxh_pair.h"

namespace oxh {

namespace {

// hydrogen-bonding energy of the pair (p, q) (DNAInteraction::
// _hydrogen_bonding, src/Interactions/DNAInteraction.cpp:779-902: zero for
// bonded neighbours), and how many of its seven factors (the Morse well
// below zero, the six angular modulations above zero) are active, the
// count the CUDA FFS backend's "nearly bonded" criterion uses
// (CUDA_DNA.cuh:1019-1030); nfac = -1 when the pair cannot bond at all
// (not complementary, or the base-base distance out of range)
__device__ __forceinline__ double hb_pair(const oxh_sys &S, const dp::Frame &fp, const dp::Frame &fq, const dv3 &rbase, int &nfac) {
	nfac = -1;
	const double rh2 = rbase.x * rbase.x + rbase.y * rbase.y + rbase.z * rbase.z;
	const int pbt = fp.btype, qbt = fq.btype;
	if(!(pbt + qbt == 3 && (double) HYDR_RCLOW * HYDR_RCLOW < rh2 && rh2 < (double) HYDR_RCHIGH * HYDR_RCHIGH)) return 0.0;
	const double hb_multi = (abs(qbt) >= 300 && abs(pbt) >= 300) ? S.hb_multi_d : 1.0;
	const double rmod = sqrt(rh2);
	const dv3 dir = {rbase.x / rmod, rbase.y / rmod, rbase.z / rmod};
	const dv3 &a1 = fp.a1, &a3 = fp.a3, &b1 = fq.a1, &b3 = fq.a3;
	const double cost1 = -dot(a1, b1), cost2 = -dot(b1, dir), cost3 = dot(a1, dir);
	const double cost4 = dot(a3, b3), cost7 = -dot(b3, dir), cost8 = dot(a3, dir);
	const double vf1 = dp::f1(S, rmod, 0, dp::type_of(qbt), dp::type_of(pbt));
	double f4[6], d;
	dp::cf4(cost1, S_HYDR_T1, f4[0], d);
	dp::cf4(cost2, S_HYDR_T2, f4[1], d);
	dp::cf4(cost3, S_HYDR_T2, f4[2], d);
	dp::cf4(cost4, S_HYDR_T4, f4[3], d);
	dp::cf4(cost7, S_HYDR_T7, f4[4], d);
	dp::cf4(cost8, S_HYDR_T7, f4[5], d);
	nfac = vf1 < 0.0 ? 1 : 0;
	for(int k = 0; k < 6; k++) nfac += f4[k] > 0.0 ? 1 : 0;
	return hb_multi * vf1 * f4[0] * f4[1] * f4[2] * f4[3] * f4[4] * f4[5];
}

// one listed pair per thread: (p, q, same strand) in the system's upload
// order, mapped to the batch's slots through slot_of
__global__ void k_pair_quantities(int n, int kind, int s, const int4 *__restrict__ pairs, const double4 *__restrict__ pos,
		const double4 *__restrict__ q64, const int2 *__restrict__ pinfo, const int2 *__restrict__ bonds, const int *__restrict__ slot_of,
		const oxh_sys *__restrict__ systems, double *__restrict__ out) {
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if(k >= n) return;
	const oxh_sys &S = systems[s];
	const int4 pr = pairs[k];
	const int p = slot_of[S.first + pr.x], q = slot_of[S.first + pr.y];
	const double4 P = pos[p], Q = pos[q];
	const bool plain = kind >= OXH_PAIR_BASE_DISTANCE && pr.z != 0;
	const dv3 r = plain ? dv3{Q.x - P.x, Q.y - P.y, Q.z - P.z} : dp::min_image_d(P, Q, S);
	Sites64 sp, sq;
	const dp::Frame fp = dp::make_frame_d(q64[p], pinfo[p].x, sp);
	const dp::Frame fq = dp::make_frame_d(q64[q], pinfo[q].x, sq);
	if(kind == OXH_PAIR_CENTRE_DISTANCE) {
		out[k] = sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
		return;
	}
	const double pb = (double) POS_BASE;
	const dv3 rbase = {(r.x + pb * sq.a1.x) - pb * sp.a1.x, (r.y + pb * sq.a1.y) - pb * sp.a1.y, (r.z + pb * sq.a1.z) - pb * sp.a1.z};
	if(kind == OXH_PAIR_BASE_DISTANCE) {
		out[k] = sqrt(rbase.x * rbase.x + rbase.y * rbase.y + rbase.z * rbase.z);
		return;
	}
	const int2 bd = bonds[p];
	int nfac = -1;
	const double e = (bd.x == q || bd.y == q) ? 0.0 : hb_pair(S, fp, fq, rbase, nfac);
	if(kind == OXH_PAIR_HB_ENERGY) out[k] = e;
	else out[k] = nfac >= 6 ? 1.0 : 0.0;   // OXH_PAIR_NEAR_HB
}

}  // namespace

void launch_pair_quantities(hipStream_t st, const Batch &b, int sys, int kind, int n, const int4 *d_pairs, double *d_out) {
	if(n <= 0) return;
	if(b.model) {
		launch_rna_pair_quantities(st, b, sys, kind, n, d_pairs, d_out);
		return;
	}
	k_pair_quantities<<<(n + 255) / 256, 256, 0, st>>>(n, kind, sys, d_pairs, b.pos, b.q64, b.pinfo, b.bonds, b.slot_of, b.d_sys, d_out);
}

}  // namespace oxh
