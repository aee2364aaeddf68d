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
// butterfly of wave shuffles, so results do not depend on timing.  PACK
// (K >= 2): workgroups of PACK_BS threads evaluate HB, CRST and CXST from
// LDS work lists (the packed form in particle_forces); same results bit
// for bit.
template<int K, bool SPLIT, bool FAR, bool PACK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_forces(int N, const double4 *__restrict__ pos, const double4 *__restrict__ q64,
		const float4 *__restrict__ q32,
		const float4 *__restrict__ bpos, const int2 *__restrict__ bonds, const int2 *__restrict__ pinfo,
		const int *__restrict__ nbr, const int2 *__restrict__ nnbr, int max_neigh, const oxh_sys *__restrict__ systems,
		oxh_sys_state *__restrict__ sstate, float4 *__restrict__ Fout, float4 *__restrict__ Tout, double *__restrict__ esplit,
		int *__restrict__ hbcount, int k, const int *__restrict__ limit, int *__restrict__ any, double *__restrict__ stress, int xcd) {
	const int blk = xcd ? xcd_block(((unsigned) N * K + blockDim.x - 1) / blockDim.x) : (int) blockIdx.x;
	if(blk < 0) return;   // whole padding workgroup
	const int t = blk * blockDim.x + threadIdx.x;
	const int i = t / K;
	const int l = t % K;
	if(k >= 0 && t == 0) *any = 0;   // inside a step: this step's list builds are done
	if constexpr(PACK) {
		// every thread of the workgroup takes part in the work lists'
		// barriers; threads past N stand in for particle N - 1, inactive
		__shared__ PackShared sh;
		const int ic = i < N ? i : N - 1;
		const int s = pinfo[ic].y;
		if(i < N && k >= 0 && l == 0 && i == systems[s].first && sstate[s].need_rebuild) {
			sstate[s].need_rebuild = 0;
			sstate[s].n_rebuilds += 1;
		}
		const bool act = i < N && !idle(i, systems[s], s, k, limit);
		particle_forces<SPLIT, FAR, true>(K, ic, l, N, pos, q64, q32, bpos, bonds, pinfo, nbr, nnbr, max_neigh, systems, sstate, Fout, Tout,
				esplit, hbcount, stress, act, &sh);
	}
	else {
		if(i >= N) return;
		const int s = pinfo[i].y;
		if(k >= 0 && l == 0 && i == systems[s].first && sstate[s].need_rebuild) {
			sstate[s].need_rebuild = 0;
			sstate[s].n_rebuilds += 1;
		}
		if(idle(i, systems[s], s, k, limit)) return;
		particle_forces<SPLIT, FAR>(K, i, l, N, pos, q64, q32, bpos, bonds, pinfo, nbr, nnbr, max_neigh, systems, sstate, Fout, Tout, esplit,
				hbcount, stress);
	}
}

// Half-list form (b.edge), first half: every bond and every close pair
// once, packed one item per lane.  Wave w takes particles [w P, w P + P):
// lane prefix sums (shuffles) over the P particles' bond flags and owned
// pair counts number the wave's items, bonds first, then the owned pairs in
// particle and row order; each lane finds its item's particle by a binary
// search over the prefix sums and evaluates it (edge_item).  Bond of
// particle i -> B1..B3[i], owned pair m of i -> E1..E3[i * ecap + m].  Same
// bookkeeping as k_forces inside a step.
template<int P>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_edge_pairs(int N, const double4 *__restrict__ pos,
		const double4 *__restrict__ q64, const float4 *__restrict__ bpos, const int2 *__restrict__ bonds, const int2 *__restrict__ pinfo,
		const int *__restrict__ eown, const int2 *__restrict__ ecount, int ecap, const oxh_sys *__restrict__ systems,
		oxh_sys_state *__restrict__ sstate, float4 *__restrict__ E1, float4 *__restrict__ E2, float4 *__restrict__ E3,
		float4 *__restrict__ B1, float4 *__restrict__ B2, float4 *__restrict__ B3, int k, int *__restrict__ any) {
	static_assert(P >= 1 && P <= 64 && (64 % P) == 0, "P must divide the wave");
	const int lane = threadIdx.x & 63;
	const int i0 = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * P;
	if(k >= 0 && blockIdx.x == 0 && threadIdx.x == 0) *any = 0;
	if(i0 >= N) return;   // whole waves
	const int ii = i0 + lane;
	int cb = 0, cp = 0;
	if(lane < P && ii < N) {
		const int s = pinfo[ii].y;
		if(k >= 0 && ii == systems[s].first && sstate[s].need_rebuild) {
			sstate[s].need_rebuild = 0;
			sstate[s].n_rebuilds += 1;
		}
		cb = bonds[ii].x >= 0;
		cp = ecount[ii].x;
	}
	int xb = cb, xp = cp;
	for(int off = 1; off < P; off <<= 1) {
		const int vb = __shfl_up(xb, off), vp = __shfl_up(xp, off);
		if(lane >= off) {
			xb += vb;
			xp += vp;
		}
	}
	const int eb = xb - cb, ep = xp - cp;   // exclusive prefix sums (lanes < P)
	const int nb = __shfl(xb, P - 1), np = __shfl(xp, P - 1);
	for(int g0 = 0; g0 < nb + np; g0 += 64) {
		const int g = g0 + lane;
		const int gp = g - nb;
		// largest particle index with prefix <= item index (wave-uniform loop)
		int lb = 0, lp = 0;
		for(int step = P >> 1; step > 0; step >>= 1) {
			if(__shfl(eb, lb + step) <= g) lb += step;
			if(__shfl(ep, lp + step) <= gp) lp += step;
		}
		const int mp = gp - __shfl(ep, lp);
		if(g < nb) {
			const int i = i0 + lb;
			const oxh_sys &S = systems[pinfo[i].y];
			int err = 0;
			float4 o1, o2, o3;
			edge_item<true>(i, bonds[i].x, pos, q64, bpos, pinfo, S, o1, o2, o3, err);
			B1[i] = o1;
			B2[i] = o2;
			B3[i] = o3;
			if(err) atomicOr(&sstate[pinfo[i].y].error, ERR_BOND);
		}
		else if(g < nb + np) {
			const int i = i0 + lp;
			const size_t slot = (size_t) i * ecap + mp;
			const oxh_sys &S = systems[pinfo[i].y];
			int err = 0;
			float4 o1, o2, o3;
			edge_item<false>(i, eown[slot], pos, q64, bpos, pinfo, S, o1, o2, o3, err);
			E1[slot] = o1;
			E2[slot] = o2;
			E3[slot] = o3;
		}
	}
}

// Half-list form, second half: particle i (K lanes) sums its side of its
// owned pairs and of its bond to n3 and the partner side of its pairs with
// j < i (slots in row ein[i]) and of its bond to n5, in a fixed order, adds
// the far pairs (Debye-Hueckel, both directions) when FAR, and writes F
// (.w energy share) and the body-frame torque (.w HB share).
template<int K, bool FAR>
__global__ void __launch_bounds__(256) k_edge_gather(int N, const double4 *__restrict__ pos, const double4 *__restrict__ q64,
		const float4 *__restrict__ q32, const float4 *__restrict__ bpos, const int2 *__restrict__ bonds, const int2 *__restrict__ pinfo,
		const int *__restrict__ nbr, const int2 *__restrict__ nnbr, int max_neigh, const oxh_sys *__restrict__ systems,
		const int *__restrict__ ein, const int2 *__restrict__ ecount, int ecap, const float4 *__restrict__ E1, const float4 *__restrict__ E2,
		const float4 *__restrict__ E3, const float4 *__restrict__ B1, const float4 *__restrict__ B2, const float4 *__restrict__ B3,
		float4 *__restrict__ Fout, float4 *__restrict__ Tout) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	const int i = t / K;
	const int l = t % K;
	if(i >= N) return;
	const int2 pi = pinfo[i];
	const oxh_sys &S = systems[pi.y];
	const int2 ec = ecount[i];
	const size_t row = (size_t) i * ecap;
	f3 F = {0, 0, 0}, T = {0, 0, 0};
	float E = 0.f, Ehb = 0.f;
	for(int m = l; m < ec.x; m += K) {
		const float4 a = E1[row + m], b = E2[row + m];
		F -= mk(a.x, a.y, a.z);
		E += a.w;
		T += mk(b.x, b.y, b.z);
		Ehb += b.w;
	}
	for(int m = l; m < ec.y; m += K) {
		const int slot = ein[row + m];
		const float4 a = E1[slot], b = E2[slot], c = E3[slot];
		F += mk(a.x, a.y, a.z);
		E += a.w;
		T += mk(c.x, c.y, c.z);
		Ehb += b.w;
	}
	const int2 bd = bonds[i];
	if(l == 0 && bd.x >= 0) {
		const float4 a = B1[i], b = B2[i];
		F -= mk(a.x, a.y, a.z);
		E += a.w;
		T += mk(b.x, b.y, b.z);
	}
	if(l == (K > 1 ? 1 : 0) && bd.y >= 0) {
		const float4 a = B1[bd.y], c = B3[bd.y];
		F += mk(a.x, a.y, a.z);
		E += a.w;
		T += mk(c.x, c.y, c.z);
	}
	Sites64 st;
	const Frame fp = make_frame_d(q64[i], 0, st);
	if(FAR) {
		const double4 Pi = pos[i];
		const float4 Bi = bpos[i];
		const int nf = nnbr[i].y;
		const int *rowf = nbr + (size_t) i * max_neigh + max_neigh - 1;
		f3 Ff = {0, 0, 0};
		float Ef = 0.f;
		for(int m = l; m < nf; m += K) {
			const int j = rowf[-m];
			const float4 Bj = bpos[j];
			f3 rb;
			if(!S.small_box) rb = min_image_f(Bi, Bj, S);
			else {
				const Frame fq = make_frame(q32[j], 0);
				rb = min_image(Pi, pos[j], S) + fq.back - fp.back;
			}
			f3 fq_ = {0, 0, 0};
			Ef += debye_huckel(S, rb, Bi.w * Bj.w, fq_);
			Ff += fq_;
		}
		F -= Ff;
		T -= cross(fp.back, Ff);
		E += 0.5f * Ef;
	}
	if(K > 1) {
		F.x = group_sum(F.x, K);
		F.y = group_sum(F.y, K);
		F.z = group_sum(F.z, K);
		T.x = group_sum(T.x, K);
		T.y = group_sum(T.y, K);
		T.z = group_sum(T.z, K);
		E = group_sum(E, K);
		Ehb = group_sum(Ehb, K);
	}
	if(l != 0) return;
	Fout[i] = make_float4(F.x, F.y, F.z, E);
	Tout[i] = make_float4(dot(fp.a1, T), dot(fp.a2, T), dot(fp.a3, T), Ehb);
}

// Far pairs on their own: Debye-Hueckel between backbone sites, few
// registers, so many more waves stay resident to hide the gather latency.
// Adds onto the F / T written by k_forces (same thread owns the
// particle's outputs in both kernels, fixed order, no atomics).  Particle
// i receives -Ff at its backbone site, i.e. the torque back x (-Ff); in the
// body frame back = (b1, b2, 0), so with Ff_body = (f1, f2, f3) that torque
// is (-b2 f3, b1 f3, b2 f1 - b1 f2).
template<int K, bool SPLIT, bool SMALL>
__global__ void __launch_bounds__(256) k_forces_far(int N, const double4 *__restrict__ pos, const float4 *__restrict__ q32,
		const float4 *__restrict__ bpos, const int *__restrict__ nbr, const int2 *__restrict__ nnbr, int max_neigh,
		const int2 *__restrict__ pinfo, const oxh_sys *__restrict__ systems, float4 *__restrict__ Fio, float4 *__restrict__ Tio,
		double *__restrict__ esplit, int k_step, const int *__restrict__ limit, double *__restrict__ stress, int xcd) {
	const int blk = xcd ? xcd_block(((unsigned) N * K + blockDim.x - 1) / blockDim.x) : (int) blockIdx.x;
	if(blk < 0) return;
	const int t = blk * blockDim.x + threadIdx.x;
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
	// SMALL: some system's box is small enough that the nearest image of two
	// backbone sites may differ from the centres' one (then the centres'
	// image plus the site offsets, as the reference takes it); compiled out
	// otherwise, which frees registers for more waves
	Frame fp;
	if(SMALL && S.small_box) fp = make_frame(q32[i], 0);
	// eight pairs per iteration, their loads issued together; each lane still
	// adds its pairs in list order
	constexpr int U = 8;
	for(int k0 = l; k0 < nf; k0 += U * K) {
		int jj[U];
		float4 Bj[U];
#pragma unroll
		for(int u = 0; u < U; u++) jj[u] = k0 + u * K < nf ? row[-(k0 + u * K)] : -1;
#pragma unroll
		for(int u = 0; u < U; u++) Bj[u] = bpos[jj[u] >= 0 ? jj[u] : i];
#pragma unroll
		for(int u = 0; u < U; u++) {
			if(jj[u] < 0) continue;
			const int j = jj[u];
			f3 rb;
			if(!(SMALL && S.small_box)) rb = min_image_f(Bi, Bj[u], S);
			else {
				const Frame fq = make_frame(q32[j], 0);
				rb = min_image(pos[i], pos[j], S) + fq.back - fp.back;
			}
			f3 fq_ = {0, 0, 0};
			Ef += debye_huckel(S, rb, Bi.w * Bj[u].w, fq_);
			Ff += fq_;
			if(SPLIT) add_virial(sv, min_image(pos[i], pos[j], S), fq_);
		}
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
// terms in FP64 (oxh::dp), F and T in FP64, far pairs inline (FAR) or in
// k_forces_far_dp.  Blocks of 64 threads; two waves per SIMD (256 VGPRs,
// a few spilled), which measured faster than one wave with every value in
// registers.
template<int K, bool SPLIT, bool FAR>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) k_forces_dp(int N, const double4 *__restrict__ pos, const double4 *__restrict__ q64,
		const float4 *__restrict__ q32, const float4 *__restrict__ bpos, const int2 *__restrict__ bonds, const int2 *__restrict__ pinfo,
		const int *__restrict__ nbr, const int2 *__restrict__ nnbr, int max_neigh, const oxh_sys *__restrict__ systems,
		oxh_sys_state *__restrict__ sstate, double4 *__restrict__ Fout, double4 *__restrict__ Tout, double *__restrict__ esplit,
		int *__restrict__ hbcount, int k, const int *__restrict__ limit, int *__restrict__ any, double *__restrict__ stress, int xcd) {
	const int blk = xcd ? xcd_block(((unsigned) N * K + blockDim.x - 1) / blockDim.x) : (int) blockIdx.x;
	if(blk < 0) return;
	const int t = blk * blockDim.x + threadIdx.x;
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
	dp::particle_forces<SPLIT, FAR>(K, i, l, N, pos, q64, q32, bpos, bonds, pinfo, nbr, nnbr, max_neigh, systems, sstate, Fout, Tout,
			esplit, hbcount, stress);
}

// All-FP64 far pairs on their own: Debye-Hueckel between FP64 backbone
// sites (centres' nearest image, then the site offsets, read from back64
// as the state mirror wrote them this step), with the few
// registers that let many waves hide the gathers, next to the
// register-heavy k_forces_dp.  Adds onto its F / T (the same thread owns a
// particle's outputs in both kernels; fixed order, no atomics).
template<int K, bool SPLIT>
__global__ void __launch_bounds__(256) k_forces_far_dp(int N, const double4 *__restrict__ pos, const double4 *__restrict__ q64,
		const double4 *__restrict__ back64, const float4 *__restrict__ bpos, const int *__restrict__ nbr, const int2 *__restrict__ nnbr, int max_neigh,
		const int2 *__restrict__ pinfo, const oxh_sys *__restrict__ systems, double4 *__restrict__ Fio, double4 *__restrict__ Tio,
		double *__restrict__ esplit, int k_step, const int *__restrict__ limit, double *__restrict__ stress, int xcd) {
	const int blk = xcd ? xcd_block(((unsigned) N * K + blockDim.x - 1) / blockDim.x) : (int) blockIdx.x;
	if(blk < 0) return;
	const int t = blk * blockDim.x + threadIdx.x;
	const int i = t / K;
	const int l = t % K;
	if(i >= N) return;
	const int s = pinfo[i].y;
	const oxh_sys &S = systems[s];
	if(idle(i, S, s, k_step, limit)) return;
	const int nf = nnbr[i].y;
	const int *row = nbr + (size_t) i * max_neigh + max_neigh - 1;
	const double4 Pi = pos[i];
	Sites64 si;
	const dp::Frame fp = dp::make_frame_d(q64[i], 0, si);
	const double cfi = bpos[i].w;
	dv3 Ff = {0, 0, 0};
	double Ef = 0.;
	double sv[6] = {0, 0, 0, 0, 0, 0};
	for(int k = l; k < nf; k += K) {
		const int j = row[-k];
		const double4 bj = back64[j];
		const dv3 rc = dp::min_image_d(Pi, pos[j], S);
		dv3 fq_ = {0, 0, 0};
		Ef += dp::debye_huckel(S, dp::site_sep(rc, dv3{bj.x, bj.y, bj.z}, si.back), cfi * (double) bpos[j].w, fq_);
		Ff += fq_;
		if(SPLIT) dp::add_virial(sv, rc, fq_);
	}
	if(K > 1) {
		Ff.x = dp::group_sum(Ff.x, K);
		Ff.y = dp::group_sum(Ff.y, K);
		Ff.z = dp::group_sum(Ff.z, K);
		Ef = dp::group_sum(Ef, K);
		if(SPLIT)
			for(int q = 0; q < 6; q++) sv[q] = dp::group_sum(sv[q], K);
	}
	if(l != 0) return;
	const dv3 Tl = -cross(fp.back, Ff);
	double4 F = Fio[i], T = Tio[i];
	F.x -= Ff.x;
	F.y -= Ff.y;
	F.z -= Ff.z;
	F.w += 0.5 * Ef;
	T.x += dot(fp.a1, Tl);
	T.y += dot(fp.a2, Tl);
	T.z += dot(fp.a3, Tl);
	Fio[i] = F;
	Tio[i] = T;
	if(SPLIT) {
		esplit[(size_t) 7 * N + i] += 0.5 * Ef;
		for(int q = 0; q < 6; q++) stress[(size_t) q * N + i] += 0.5 * sv[q];
	}
}

template<int K, bool FAR>
static void launch_dp_close(hipStream_t st, const Batch &b, bool split, int k) {
	const int block = 64;
	size_t threads = (size_t) b.N * K;
	const unsigned nblk = (unsigned) ((threads + block - 1) / block);
	dim3 grid(b.xcd ? xcd_grid(nblk) : nblk);
	const int *lim = k >= 0 ? b.limit : nullptr;
	if(split)
		k_forces_dp<K, true, FAR><<<grid, block, 0, st>>>(b.N, b.pos, b.q64, b.q32, b.bpos, b.bonds, b.pinfo, b.nbr, b.nnbr, b.max_neigh,
				b.d_sys, b.d_state, b.F64, b.T64, b.esplit, b.hbcount, k, lim, b.any_rebuild, b.stress, b.xcd);
	else
		k_forces_dp<K, false, FAR><<<grid, block, 0, st>>>(b.N, b.pos, b.q64, b.q32, b.bpos, b.bonds, b.pinfo, b.nbr, b.nnbr, b.max_neigh,
				b.d_sys, b.d_state, b.F64, b.T64, nullptr, b.want_hbcount ? b.hbcount : nullptr, k, lim, b.any_rebuild, nullptr, b.xcd);
}

template<int K>
static void launch_far_dp(hipStream_t st, const Batch &b, bool split, int k) {
	const int block = 256;
	size_t threads = (size_t) b.N * K;
	const unsigned nblk = (unsigned) ((threads + block - 1) / block);
	dim3 grid(b.xcd ? xcd_grid(nblk) : nblk);
	const int *lim = k >= 0 ? b.limit : nullptr;
	if(split)
		k_forces_far_dp<K, true><<<grid, block, 0, st>>>(b.N, b.pos, b.q64, b.back64, b.bpos, b.nbr, b.nnbr, b.max_neigh, b.pinfo, b.d_sys, b.F64,
				b.T64, b.esplit, k, lim, b.stress, b.xcd);
	else
		k_forces_far_dp<K, false><<<grid, block, 0, st>>>(b.N, b.pos, b.q64, b.back64, b.bpos, b.nbr, b.nnbr, b.max_neigh, b.pinfo, b.d_sys, b.F64,
				b.T64, nullptr, k, lim, nullptr, b.xcd);
}

// all-FP64 forces: close (+ bonded) kernel with b.lanes, far pairs inline
// (b.far_lanes == 0) or in k_forces_far_dp with b.far_lanes
template<int K>
static void launch_dp(hipStream_t st, const Batch &b, bool split, int k) {
	if(b.far_lanes == 0) {
		launch_dp_close<K, true>(st, b, split, k);
		return;
	}
	launch_dp_close<K, false>(st, b, split, k);
	switch(b.far_lanes) {
	case 4: launch_far_dp<4>(st, b, split, k); break;
	case 16: launch_far_dp<16>(st, b, split, k); break;
	case 32: launch_far_dp<32>(st, b, split, k); break;
	default: launch_far_dp<8>(st, b, split, k); break;
	}
}

template<int K, bool PACK>
static void launch_kp(hipStream_t st, const Batch &b, bool split, int block, bool far_inline, int k) {
	size_t threads = (size_t) b.N * K;
	const unsigned nblk = (unsigned) ((threads + block - 1) / block);
	dim3 grid(b.xcd ? xcd_grid(nblk) : nblk);
	int *hbc = b.want_hbcount ? b.hbcount : nullptr;
	if(split) {
		if(far_inline)
			k_forces<K, true, true, PACK><<<grid, block, 0, st>>>(b.N, b.pos, b.q64, b.q32, b.bpos, b.bonds, b.pinfo, b.nbr, b.nnbr,
					b.max_neigh, b.d_sys, b.d_state, b.F, b.T, b.esplit, b.hbcount, k, k >= 0 ? b.limit : nullptr, b.any_rebuild, b.stress, b.xcd);
		else
			k_forces<K, true, false, PACK><<<grid, block, 0, st>>>(b.N, b.pos, b.q64, b.q32, b.bpos, b.bonds, b.pinfo, b.nbr, b.nnbr,
					b.max_neigh, b.d_sys, b.d_state, b.F, b.T, b.esplit, b.hbcount, k, k >= 0 ? b.limit : nullptr, b.any_rebuild, b.stress, b.xcd);
	}
	else {
		if(far_inline)
			k_forces<K, false, true, PACK><<<grid, block, 0, st>>>(b.N, b.pos, b.q64, b.q32, b.bpos, b.bonds, b.pinfo, b.nbr, b.nnbr,
					b.max_neigh, b.d_sys, b.d_state, b.F, b.T, nullptr, hbc, k, k >= 0 ? b.limit : nullptr, b.any_rebuild, nullptr, b.xcd);
		else
			k_forces<K, false, false, PACK><<<grid, block, 0, st>>>(b.N, b.pos, b.q64, b.q32, b.bpos, b.bonds, b.pinfo, b.nbr, b.nnbr,
					b.max_neigh, b.d_sys, b.d_state, b.F, b.T, nullptr, hbc, k, k >= 0 ? b.limit : nullptr, b.any_rebuild, nullptr, b.xcd);
	}
}

// b.pack: the packed form, in workgroups of PACK_BS threads
template<int K>
static void launch_k(hipStream_t st, const Batch &b, bool split, int block, bool far_inline, int k) {
	if(b.pack) launch_kp<K, true>(st, b, split, PACK_BS, far_inline, k);
	else launch_kp<K, false>(st, b, split, block, far_inline, k);
}

template<int K, bool SMALL>
static void launch_far_k(hipStream_t st, const Batch &b, bool split, int block, int k) {
	size_t threads = (size_t) b.N * K;
	const unsigned nblk = (unsigned) ((threads + block - 1) / block);
	dim3 grid(b.xcd ? xcd_grid(nblk) : nblk);
	if(split)
		k_forces_far<K, true, SMALL><<<grid, block, 0, st>>>(b.N, b.pos, b.q32, b.bpos, b.nbr, b.nnbr, b.max_neigh, b.pinfo, b.d_sys, b.F,
				b.T, b.esplit, k, k >= 0 ? b.limit : nullptr, b.stress, b.xcd);
	else
		k_forces_far<K, false, SMALL><<<grid, block, 0, st>>>(b.N, b.pos, b.q32, b.bpos, b.nbr, b.nnbr, b.max_neigh, b.pinfo, b.d_sys, b.F,
				b.T, nullptr, k, k >= 0 ? b.limit : nullptr, nullptr, b.xcd);
}

// slot pools take any system later, so they keep the small-box path
template<int K>
static void launch_far(hipStream_t st, const Batch &b, bool split, int block, int k) {
	bool small = b.limit != nullptr;
	for(const oxh_sys &S : b.h_sys) small |= S.small_box != 0;
	if(small) launch_far_k<K, true>(st, b, split, block, k);
	else launch_far_k<K, false>(st, b, split, block, k);
}

template<int P>
static void launch_edge_pairs(hipStream_t st, const Batch &b, int k) {
	const size_t waves = ((size_t) b.N + P - 1) / P;
	dim3 grid((unsigned) ((waves + 3) / 4));
	k_edge_pairs<P><<<grid, 256, 0, st>>>(b.N, b.pos, b.q64, b.bpos, b.bonds, b.pinfo, b.eown, b.ecount, b.ecap, b.d_sys, b.d_state, b.E1,
			b.E2, b.E3, b.B1, b.B2, b.B3, k, b.any_rebuild);
}

template<int K>
static void launch_edge_gather(hipStream_t st, const Batch &b, bool far_inline) {
	dim3 grid((unsigned) (((size_t) b.N * K + 255) / 256));
	if(far_inline)
		k_edge_gather<K, true><<<grid, 256, 0, st>>>(b.N, b.pos, b.q64, b.q32, b.bpos, b.bonds, b.pinfo, b.nbr, b.nnbr, b.max_neigh, b.d_sys,
				b.ein, b.ecount, b.ecap, b.E1, b.E2, b.E3, b.B1, b.B2, b.B3, b.F, b.T);
	else
		k_edge_gather<K, false><<<grid, 256, 0, st>>>(b.N, b.pos, b.q64, b.q32, b.bpos, b.bonds, b.pinfo, b.nbr, b.nnbr, b.max_neigh, b.d_sys,
				b.ein, b.ecount, b.ecap, b.E1, b.E2, b.E3, b.B1, b.B2, b.B3, b.F, b.T);
}

// half-list form: b.edge_group particles per wave in the pair kernel,
// b.lanes lanes per particle in the gather
static void launch_edge(hipStream_t st, const Batch &b, bool far_inline, int k) {
	switch(b.edge_group) {
	case 4: launch_edge_pairs<4>(st, b, k); break;
	case 8: launch_edge_pairs<8>(st, b, k); break;
	case 32: launch_edge_pairs<32>(st, b, k); break;
	case 64: launch_edge_pairs<64>(st, b, k); break;
	default: launch_edge_pairs<16>(st, b, k); break;
	}
	switch(b.lanes) {
	case 1: launch_edge_gather<1>(st, b, far_inline); break;
	case 2: launch_edge_gather<2>(st, b, far_inline); break;
	case 4: launch_edge_gather<4>(st, b, far_inline); break;
	case 16: launch_edge_gather<16>(st, b, far_inline); break;
	case 32: launch_edge_gather<32>(st, b, far_inline); break;
	default: launch_edge_gather<8>(st, b, far_inline); break;
	}
}

// close (+ bonded) kernel; the far pairs either inline (far_lanes == 0) or
// in k_forces_far right after it.  The step's forces (split = false) take
// the half-list form when b.edge is set.
namespace {
void launch_pair_forces(hipStream_t st, const Batch &b, bool split, int k) {
	if(b.N == 0) return;
	if(b.model) {
		launch_rna_forces(st, b, split, k);
		return;
	}
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
	if(b.edge && !split) launch_edge(st, b, inl, k);
	else {
		switch(b.lanes) {
		case 1: launch_k<1>(st, b, split, block, inl, k); break;
		case 2: launch_k<2>(st, b, split, block, inl, k); break;
		case 4: launch_k<4>(st, b, split, block, inl, k); break;
		case 8: launch_k<8>(st, b, split, block, inl, k); break;
		case 16: launch_k<16>(st, b, split, block, inl, k); break;
		case 32: launch_k<32>(st, b, split, block, inl, k); break;
		default: launch_k<64>(st, b, split, block, inl, k); break;
		}
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
}  // namespace

// pair forces (written, not accumulated), then the external forces added
// on top
void launch_forces(hipStream_t st, const Batch &b, bool split, int k) {
	launch_pair_forces(st, b, split, k);
	launch_ext_forces(st, b);
}

}  // namespace oxh
