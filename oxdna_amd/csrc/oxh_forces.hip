// oxDNA2 forces and torques on gfx950.
//
// K threads per nucleotide (K picked by timing at upload) over a full
// (both-direction) neighbour list split into a close sublist (every term)
// and a far sublist (Debye-Hueckel only); each nucleotide's K partial sums
// are combined by a fixed shuffle butterfly, so there are no atomics and
// results are bitwise reproducible.  Nearest-image separations are formed
// in FP64 from FP64 positions and rounded to FP32 once; all pair terms are
// FP32.
//
// Terms follow the reference's CPU DNA2_nomesh formulation
// (src/Interactions/DNAInteraction.cpp:415-1205,
//  src/Interactions/DNA2Interaction.cpp:157-306); the kernel replaces the
// reference's dna_forces (src/CUDA/Interactions/CUDA_DNA.cuh:832-904).
// Output contract as CUDABaseInteraction::compute_forces
// (src/CUDA/Interactions/CUDABaseInteraction.h:60): F.xyz lab-frame force,
// T.xyz body-frame torque; F.w = this particle's share of the potential
// energy (half of each pair, so that the sum over particles is U), T.w =
// its share of the hydrogen-bonding energy.
#include "oxh_device.h"
#include "oxh_internal.h"
#include "oxh_pair.h"

namespace oxh {

// K threads per nucleotide: lane l of nucleotide i takes list entries
// l, l + K, ... of i's close sublist, then of its far sublist, and the n3 /
// n5 bonded pairs go to lanes 0 / 1.  Partial sums are combined by a fixed
// butterfly of wave shuffles, so results do not depend on timing.
template<int K, bool SPLIT, bool FAR>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_forces(int N, const double4 *__restrict__ pos, const double4 *__restrict__ q64,
		const float4 *__restrict__ q32,
		const float4 *__restrict__ bpos, const int2 *__restrict__ bonds, const int2 *__restrict__ pinfo,
		const int *__restrict__ nbr, const int2 *__restrict__ nnbr, int max_neigh, const oxh_sys *__restrict__ systems,
		oxh_sys_state *__restrict__ sstate, float4 *__restrict__ Fout, float4 *__restrict__ Tout, double *__restrict__ esplit,
		int *__restrict__ hbcount, int k, const int *__restrict__ limit, int *__restrict__ any, double *__restrict__ stress) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	const int i = t / K;
	const int l = t % K;
	if(k >= 0 && t == 0) *any = 0;   // inside a step: this step's list builds are done
	if(i >= N) return;
	const int s = pinfo[i].y;
	if(k >= 0 && l == 0 && i == systems[s].first && sstate[s].need_rebuild) {
		sstate[s].need_rebuild = 0;
		sstate[s].n_rebuilds += 1;
	}
	if(idle(i, systems[s], s, k, limit)) return;

	particle_forces<SPLIT, FAR>(K, i, l, N, pos, q64, q32, bpos, bonds, pinfo, nbr, nnbr, max_neigh, systems, sstate, Fout, Tout, esplit, hbcount,
			stress);
}

// Far pairs on their own: Debye-Hueckel between backbone sites, few
// registers, so many more waves stay resident to hide the gather latency.
// Adds onto the F / T written by k_forces (same thread owns the
// particle's outputs in both kernels, fixed order, no atomics).  Particle
// i receives -Ff at its backbone site, i.e. the torque back x (-Ff); in the
// body frame back = (b1, b2, 0), so with Ff_body = (f1, f2, f3) that torque
// is (-b2 f3, b1 f3, b2 f1 - b1 f2).
template<int K, bool SPLIT>
__global__ void __launch_bounds__(256) k_forces_far(int N, const double4 *__restrict__ pos, const float4 *__restrict__ q32,
		const float4 *__restrict__ bpos, const int *__restrict__ nbr, const int2 *__restrict__ nnbr, int max_neigh,
		const int2 *__restrict__ pinfo, const oxh_sys *__restrict__ systems, float4 *__restrict__ Fio, float4 *__restrict__ Tio,
		double *__restrict__ esplit, int k_step, const int *__restrict__ limit, double *__restrict__ stress) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	const int i = t / K;
	const int l = t % K;
	if(i >= N) return;
	const int s = pinfo[i].y;
	const oxh_sys &S = systems[s];
	if(idle(i, S, s, k_step, limit)) return;
	const int nf = nnbr[i].y;
	const int *row = nbr + (size_t) i * max_neigh + max_neigh - 1;
	const float4 Bi = bpos[i];
	f3 Ff = {0, 0, 0};
	float Ef = 0.f;
	float sv[6] = {0, 0, 0, 0, 0, 0};
	Frame fp;
	if(S.small_box) fp = make_frame(q32[i], 0);
	for(int k = l; k < nf; k += K) {
		const int j = row[-k];
		const float4 Bj = bpos[j];
		f3 rb;
		if(!S.small_box) rb = min_image_f(Bi, Bj, S);
		else {
			const Frame fq = make_frame(q32[j], 0);
			rb = min_image(pos[i], pos[j], S) + fq.back - fp.back;
		}
		f3 fq_ = {0, 0, 0};
		Ef += debye_huckel(S, rb, Bi.w * Bj.w, fq_);
		Ff += fq_;
		if(SPLIT) add_virial(sv, min_image(pos[i], pos[j], S), fq_);
	}
	if(K > 1) {
		Ff.x = group_sum(Ff.x, K);
		Ff.y = group_sum(Ff.y, K);
		Ff.z = group_sum(Ff.z, K);
		Ef = group_sum(Ef, K);
		if(SPLIT)
			for(int q = 0; q < 6; q++) sv[q] = group_sum(sv[q], K);
	}
	if(l != 0) return;
	f3 a1, a2, a3;
	quat_axes(q32[i], a1, a2, a3);
	const float f1b = dot(a1, Ff), f2b = dot(a2, Ff), f3b = dot(a3, Ff);
	float4 F = Fio[i], T = Tio[i];
	F.x -= Ff.x;
	F.y -= Ff.y;
	F.z -= Ff.z;
	F.w += 0.5f * Ef;
	T.x += -POS_MM_BACK2 * f3b;
	T.y += POS_MM_BACK1 * f3b;
	T.z += POS_MM_BACK2 * f1b - POS_MM_BACK1 * f2b;
	Fio[i] = F;
	Tio[i] = T;
	if(SPLIT) {
		esplit[(size_t) 7 * N + i] += 0.5f * Ef;
		for(int q = 0; q < 6; q++) stress[(size_t) q * N + i] += 0.5f * sv[q];
	}
}

// All-FP64 forces (the reference's backend_precision = double): the same
// terms in FP64 (oxh::dp), far pairs inline, F and T in FP64.  Blocks of
// 64 threads, so the FP64 terms get the register file they need.
template<int K, bool SPLIT>
__global__ void __launch_bounds__(64) k_forces_dp(int N, const double4 *__restrict__ pos, const double4 *__restrict__ q64,
		const float4 *__restrict__ q32, const float4 *__restrict__ bpos, const int2 *__restrict__ bonds, const int2 *__restrict__ pinfo,
		const int *__restrict__ nbr, const int2 *__restrict__ nnbr, int max_neigh, const oxh_sys *__restrict__ systems,
		oxh_sys_state *__restrict__ sstate, double4 *__restrict__ Fout, double4 *__restrict__ Tout, double *__restrict__ esplit,
		int *__restrict__ hbcount, int k, const int *__restrict__ limit, int *__restrict__ any, double *__restrict__ stress) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	const int i = t / K;
	const int l = t % K;
	if(k >= 0 && t == 0) *any = 0;
	if(i >= N) return;
	const int s = pinfo[i].y;
	if(k >= 0 && l == 0 && i == systems[s].first && sstate[s].need_rebuild) {
		sstate[s].need_rebuild = 0;
		sstate[s].n_rebuilds += 1;
	}
	if(idle(i, systems[s], s, k, limit)) return;
	dp::particle_forces<SPLIT, true>(K, i, l, N, pos, q64, q32, bpos, bonds, pinfo, nbr, nnbr, max_neigh, systems, sstate, Fout, Tout,
			esplit, hbcount, stress);
}

template<int K>
static void launch_dp(hipStream_t st, const Batch &b, bool split, int k) {
	const int block = 64;
	size_t threads = (size_t) b.N * K;
	dim3 grid((unsigned) ((threads + block - 1) / block));
	const int *lim = k >= 0 ? b.limit : nullptr;
	if(split)
		k_forces_dp<K, true><<<grid, block, 0, st>>>(b.N, b.pos, b.q64, b.q32, b.bpos, b.bonds, b.pinfo, b.nbr, b.nnbr, b.max_neigh, b.d_sys,
				b.d_state, b.F64, b.T64, b.esplit, b.hbcount, k, lim, b.any_rebuild, b.stress);
	else
		k_forces_dp<K, false><<<grid, block, 0, st>>>(b.N, b.pos, b.q64, b.q32, b.bpos, b.bonds, b.pinfo, b.nbr, b.nnbr, b.max_neigh,
				b.d_sys, b.d_state, b.F64, b.T64, nullptr, b.want_hbcount ? b.hbcount : nullptr, k, lim, b.any_rebuild, nullptr);
}

template<int K>
static void launch_k(hipStream_t st, const Batch &b, bool split, int block, bool far_inline, int k) {
	size_t threads = (size_t) b.N * K;
	dim3 grid((unsigned) ((threads + block - 1) / block));
	int *hbc = b.want_hbcount ? b.hbcount : nullptr;
	if(split) {
		if(far_inline)
			k_forces<K, true, true><<<grid, block, 0, st>>>(b.N, b.pos, b.q64, b.q32, b.bpos, b.bonds, b.pinfo, b.nbr, b.nnbr, b.max_neigh,
					b.d_sys, b.d_state, b.F, b.T, b.esplit, b.hbcount, k, k >= 0 ? b.limit : nullptr, b.any_rebuild, b.stress);
		else
			k_forces<K, true, false><<<grid, block, 0, st>>>(b.N, b.pos, b.q64, b.q32, b.bpos, b.bonds, b.pinfo, b.nbr, b.nnbr, b.max_neigh,
					b.d_sys, b.d_state, b.F, b.T, b.esplit, b.hbcount, k, k >= 0 ? b.limit : nullptr, b.any_rebuild, b.stress);
	}
	else {
		if(far_inline)
			k_forces<K, false, true><<<grid, block, 0, st>>>(b.N, b.pos, b.q64, b.q32, b.bpos, b.bonds, b.pinfo, b.nbr, b.nnbr, b.max_neigh,
					b.d_sys, b.d_state, b.F, b.T, nullptr, hbc, k, k >= 0 ? b.limit : nullptr, b.any_rebuild, nullptr);
		else
			k_forces<K, false, false><<<grid, block, 0, st>>>(b.N, b.pos, b.q64, b.q32, b.bpos, b.bonds, b.pinfo, b.nbr, b.nnbr, b.max_neigh,
					b.d_sys, b.d_state, b.F, b.T, nullptr, hbc, k, k >= 0 ? b.limit : nullptr, b.any_rebuild, nullptr);
	}
}

template<int K>
static void launch_far(hipStream_t st, const Batch &b, bool split, int block, int k) {
	size_t threads = (size_t) b.N * K;
	dim3 grid((unsigned) ((threads + block - 1) / block));
	if(split)
		k_forces_far<K, true><<<grid, block, 0, st>>>(b.N, b.pos, b.q32, b.bpos, b.nbr, b.nnbr, b.max_neigh, b.pinfo, b.d_sys, b.F, b.T,
				b.esplit, k, k >= 0 ? b.limit : nullptr, b.stress);
	else
		k_forces_far<K, false><<<grid, block, 0, st>>>(b.N, b.pos, b.q32, b.bpos, b.nbr, b.nnbr, b.max_neigh, b.pinfo, b.d_sys, b.F, b.T,
				nullptr, k, k >= 0 ? b.limit : nullptr, nullptr);
}

// close (+ bonded) kernel; the far pairs either inline (far_lanes == 0) or
// in k_forces_far right after it
void launch_forces(hipStream_t st, const Batch &b, bool split, int k) {
	if(b.N == 0) return;
	if(b.fp64) {
		switch(b.lanes) {
		case 1: launch_dp<1>(st, b, split, k); break;
		case 2: launch_dp<2>(st, b, split, k); break;
		case 4: launch_dp<4>(st, b, split, k); break;
		case 16: launch_dp<16>(st, b, split, k); break;
		case 32: launch_dp<32>(st, b, split, k); break;
		default: launch_dp<8>(st, b, split, k); break;
		}
		return;
	}
	const bool inl = b.far_lanes == 0;
	int block = b.force_block;
	switch(b.lanes) {
	case 1: launch_k<1>(st, b, split, block, inl, k); break;
	case 2: launch_k<2>(st, b, split, block, inl, k); break;
	case 4: launch_k<4>(st, b, split, block, inl, k); break;
	case 8: launch_k<8>(st, b, split, block, inl, k); break;
	case 16: launch_k<16>(st, b, split, block, inl, k); break;
	case 32: launch_k<32>(st, b, split, block, inl, k); break;
	default: launch_k<64>(st, b, split, block, inl, k); break;
	}
	if(inl) return;
	block = b.far_block;
	switch(b.far_lanes) {
	case 1: launch_far<1>(st, b, split, block, k); break;
	case 2: launch_far<2>(st, b, split, block, k); break;
	case 4: launch_far<4>(st, b, split, block, k); break;
	case 8: launch_far<8>(st, b, split, block, k); break;
	case 16: launch_far<16>(st, b, split, block, k); break;
	case 32: launch_far<32>(st, b, split, block, k); break;
	default: launch_far<64>(st, b, split, block, k); break;
	}
}

}  // namespace oxh
