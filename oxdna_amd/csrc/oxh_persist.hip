// Workgroup-resident MD engine for small systems on gfx950.
//
// One workgroup integrates one system for a whole chunk of steps in a
// single launch: the phases of a step (first half kick + drift + rotation,
// skin check, Verlet rebuild when due, forces, second half kick +
// thermostat, Bussi rescaling) are separated by workgroup barriers instead
// of kernel launches.  A system of a few dozen nucleotides costs microseconds
// per step on one CU, so a batch of independent small systems (hairpins,
// short duplexes, short served jobs) runs one system per CU with no launch
// or grid-synchronisation cost per step.
//
// Every phase calls the same per-particle device functions as the
// one-kernel-per-phase path (oxh_step.h, oxh_list.h, oxh_pair.h) on the
// same HBM arrays, so with the same lanes per nucleotide both paths give the
// same bits.  The lanes per nucleotide depend only on the system's size,
// so a system's trajectory does not depend on the other systems of the
// batch.
//
// Reference semantics: the per-step sequence of MD_CUDABackend::sim_step
// (src/CUDA/Backends/MD_CUDABackend.cu:567-619) with the mixed-precision
// kernels (src/CUDA/Backends/CUDA_mixed.cuh:23-93), the simple Verlet list
// (src/CUDA/Lists/CUDA_simple_verlet.cuh:40-82) and the thermostats
// (src/CUDA/Thermostats/CUDA{Langevin,Bussi,Brownian}Thermostat.cu).
#include "oxh_device.h"
#include "oxh_internal.h"
#include "oxh_list.h"
#include "oxh_pair.h"
#include "oxh_step.h"

namespace oxh {

namespace {

struct PersistArgs {
	int N, max_neigh, nsteps, cell_cap;
	double4 *pos, *vel, *Lm, *q64, *rlist;
	float4 *q32, *F, *T, *bpos;
	const int2 *bonds, *pinfo;
	const int *orig;
	long long *pstep;
	int *nbr;
	int2 *nnbr;
	const oxh_sys *systems;
	oxh_sys_state *state;
	const int *cell_count, *cell_items;
	const float4 *pos32;
	int *hbcount;
	const int *limit;
};

template<int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_persist(PersistArgs a) {
	__shared__ double sh[5][256];
	__shared__ double fac[5];
	__shared__ int need;
	const int s = blockIdx.x;
	const oxh_sys &S = a.systems[s];
	const int first = S.first, n = S.n;
	const int nsteps = a.limit ? min(a.nsteps, a.limit[s]) : a.nsteps;
	if(n == 0 || nsteps <= 0) return;   // empty slot, or no steps left for this system
	const int tid = threadIdx.x;
	const int K = persist_lanes(n, BLOCK);
	const long long step0 = a.pstep[first];
	if(tid == 0) need = 0;
	__syncthreads();
	for(int it = 0; it < nsteps; it++) {
		// 1. first half kick, drift, rotation, FP32 mirror, skin check
		if(tid < n && first_step_particle(first + tid, S, a.pos, a.vel, a.Lm, a.q64, a.q32, a.bpos, nullptr, a.F, a.T, a.rlist, a.bonds, a.pstep, nullptr))
			need = 1;
		__syncthreads();
		// 2. Verlet rebuild (all pairs of the system; such systems use no cells)
		const bool rebuild = need != 0;
		if(rebuild) {
			for(int u = tid; u < n * BUILD_LANES; u += BLOCK)
				build_row<BUILD_LANES>(first + u / BUILD_LANES, u % BUILD_LANES, s, a.pos, a.bonds, a.systems, a.state, a.cell_count,
						a.cell_items, a.cell_cap, a.nbr, a.nnbr, a.max_neigh, a.rlist, a.pos32, nullptr);
			__syncthreads();
		}
		// 3. forces and torques, K lanes per nucleotide, far pairs inline
		if(tid < n * K)
			particle_forces<false, true>(K, first + tid / K, tid % K, a.N, a.pos, a.q64, a.q32, a.bpos, a.bonds, a.pinfo, a.nbr, a.nnbr,
					a.max_neigh, a.systems, a.state, a.F, a.T, nullptr, a.hbcount, nullptr);
		__syncthreads();
		// 4. second half kick + Langevin / Brownian
		if(tid < n) second_step_particle(first + tid, S, a.vel, a.Lm, a.F, a.T, a.orig, a.pstep);
		if(tid == 0) {
			if(rebuild) a.state[s].n_rebuilds += 1;
			need = 0;
		}
		// 5. Bussi: kinetic energies about the centre of mass, new targets,
		// rescaling
		const long long step = step0 + it;
		if(S.thermostat == 2 && step % S.newtonian_steps == 0) {
			__syncthreads();
			bussi_chunk_sums(first, first + n, a.vel, a.Lm, sh);
			if(tid == 0) {
				double r[5];
				for(int k = 0; k < 5; k++) r[k] = 0.0 + sh[k][0];
				double ft, fr, vc[3];
				bussi_update(S, a.state[s], step, r, ft, fr, vc);
				fac[0] = ft;
				fac[1] = fr;
				fac[2] = vc[0];
				fac[3] = vc[1];
				fac[4] = vc[2];
			}
			__syncthreads();
			if(tid < n) bussi_rescale_particle(first + tid, a.vel, a.Lm, fac[0], fac[1], fac[2], fac[3], fac[4]);
		}
		__syncthreads();
	}
}

// The same engine with the system held in LDS for the whole launch
// (systems of at most PERSIST_LDS_MAX_N nucleotides): the state, the FP32
// mirror, forces, topology, RNG counters and the Verlet list (row stride
// n - 1, which no row can exceed) are staged in at the start and written
// back at the end, with particle indices local to the system (a copy of
// its parameters with first = 0).  Every gather of a step then comes from
// LDS instead of L2.  The arithmetic and its order are those of k_persist,
// so both give the same bits.
struct LdsLayout {
	double4 *pos, *vel, *Lm, *q64, *rlist;
	float4 *q32, *F, *T, *bpos;
	int2 *bonds, *pinfo, *nnbr;
	long long *pstep;
	int *orig, *nbr;
};

__device__ __forceinline__ LdsLayout lds_layout(unsigned char *base, int nmax) {
	LdsLayout L;
	L.pos = (double4 *) base;
	L.vel = L.pos + nmax;
	L.Lm = L.vel + nmax;
	L.q64 = L.Lm + nmax;
	L.rlist = L.q64 + nmax;
	L.q32 = (float4 *) (L.rlist + nmax);
	L.F = L.q32 + nmax;
	L.T = L.F + nmax;
	L.bpos = L.T + nmax;
	L.pstep = (long long *) (L.bpos + nmax);
	L.bonds = (int2 *) (L.pstep + nmax);
	L.pinfo = L.bonds + nmax;
	L.nnbr = L.pinfo + nmax;
	L.orig = (int *) (L.nnbr + nmax);
	L.nbr = L.orig + nmax;
	return L;
}

template<int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_persist_lds(PersistArgs a, int nmax) {
	extern __shared__ __align__(32) unsigned char lds[];
	__shared__ double sh[5][256];
	__shared__ double fac[5];
	__shared__ int need;
	__shared__ oxh_sys S;
	const int s = blockIdx.x;
	const oxh_sys &G = a.systems[s];
	const int first = G.first, n = G.n;
	const int nsteps = a.limit ? min(a.nsteps, a.limit[s]) : a.nsteps;
	if(n == 0 || nsteps <= 0) return;
	const int tid = threadIdx.x;
	const int K = persist_lanes(n, BLOCK);
	const int stride = n > 1 ? n - 1 : 1;
	const LdsLayout L = lds_layout(lds, nmax);
	oxh_sys_state *st = a.state + s;
	int *hbc = a.hbcount ? a.hbcount + first : nullptr;
	// stage in, indices made local to the system
	if(tid == 0) {
		S = G;
		S.first = 0;
		need = 0;
	}
	for(int i = tid; i < n; i += BLOCK) {
		const int g = first + i;
		L.pos[i] = a.pos[g];
		L.vel[i] = a.vel[g];
		L.Lm[i] = a.Lm[g];
		L.q64[i] = a.q64[g];
		L.rlist[i] = a.rlist[g];
		L.q32[i] = a.q32[g];
		L.F[i] = a.F[g];
		L.T[i] = a.T[g];
		L.bpos[i] = a.bpos[g];
		L.pstep[i] = a.pstep[g];
		const int2 bd = a.bonds[g];
		L.bonds[i] = make_int2(bd.x >= 0 ? bd.x - first : -1, bd.y >= 0 ? bd.y - first : -1);
		L.pinfo[i] = make_int2(a.pinfo[g].x, 0);
		L.orig[i] = a.orig[g];
		const int2 nn = a.nnbr[g];
		L.nnbr[i] = nn;
		const int *grow = a.nbr + (size_t) g * a.max_neigh;
		int *lrow = L.nbr + (size_t) i * stride;
		for(int k = 0; k < nn.x; k++) lrow[k] = grow[k] - first;
		for(int k = 0; k < nn.y; k++) lrow[stride - 1 - k] = grow[a.max_neigh - 1 - k] - first;
	}
	const long long step0 = a.pstep[first];
	__syncthreads();
	for(int it = 0; it < nsteps; it++) {
		// 1. first half kick, drift, rotation, FP32 mirror, skin check
		if(tid < n && first_step_particle(tid, S, L.pos, L.vel, L.Lm, L.q64, L.q32, L.bpos, nullptr, L.F, L.T, L.rlist, L.bonds, L.pstep, nullptr))
			need = 1;
		__syncthreads();
		// 2. Verlet rebuild (all pairs of the system)
		const bool rebuild = need != 0;
		if(rebuild) {
			for(int u = tid; u < n * BUILD_LANES; u += BLOCK)
				build_row<BUILD_LANES>(u / BUILD_LANES, u % BUILD_LANES, 0, L.pos, L.bonds, &S, st, a.cell_count, a.cell_items,
						a.cell_cap, L.nbr, L.nnbr, stride, L.rlist, a.pos32, nullptr);
			__syncthreads();
		}
		// 3. forces and torques, K lanes per nucleotide, far pairs inline
		if(tid < n * K)
			particle_forces<false, true>(K, tid / K, tid % K, n, L.pos, L.q64, L.q32, L.bpos, L.bonds, L.pinfo, L.nbr, L.nnbr, stride, &S,
					st, L.F, L.T, nullptr, hbc, nullptr);
		__syncthreads();
		// 4. second half kick + Langevin / Brownian
		if(tid < n) second_step_particle(tid, S, L.vel, L.Lm, L.F, L.T, L.orig, L.pstep);
		if(tid == 0) {
			if(rebuild) st->n_rebuilds += 1;
			need = 0;
		}
		// 5. Bussi
		const long long step = step0 + it;
		if(S.thermostat == 2 && step % S.newtonian_steps == 0) {
			__syncthreads();
			bussi_chunk_sums(0, n, L.vel, L.Lm, sh);
			if(tid == 0) {
				double r[5];
				for(int k = 0; k < 5; k++) r[k] = 0.0 + sh[k][0];
				double ft, fr, vc[3];
				bussi_update(S, *st, step, r, ft, fr, vc);
				fac[0] = ft;
				fac[1] = fr;
				fac[2] = vc[0];
				fac[3] = vc[1];
				fac[4] = vc[2];
			}
			__syncthreads();
			if(tid < n) bussi_rescale_particle(tid, L.vel, L.Lm, fac[0], fac[1], fac[2], fac[3], fac[4]);
		}
		__syncthreads();
	}
	// stage out (the list with the batch's indices and row stride; a row
	// longer than the batch's rows is reported as the grid build reports
	// it, and the host replays the chunk with larger rows)
	for(int i = tid; i < n; i += BLOCK) {
		const int g = first + i;
		a.pos[g] = L.pos[i];
		a.vel[g] = L.vel[i];
		a.Lm[g] = L.Lm[i];
		a.q64[g] = L.q64[i];
		a.rlist[g] = L.rlist[i];
		a.q32[g] = L.q32[i];
		a.F[g] = L.F[i];
		a.T[g] = L.T[i];
		a.bpos[g] = L.bpos[i];
		a.pstep[g] = L.pstep[i];
		int2 nn = L.nnbr[i];
		if(nn.x + nn.y > a.max_neigh) {
			atomicOr(&st->error, ERR_LIST);
			nn = make_int2(0, 0);
		}
		a.nnbr[g] = nn;
		int *grow = a.nbr + (size_t) g * a.max_neigh;
		const int *lrow = L.nbr + (size_t) i * stride;
		for(int k = 0; k < nn.x; k++) grow[k] = lrow[k] + first;
		for(int k = 0; k < nn.y; k++) grow[a.max_neigh - 1 - k] = lrow[stride - 1 - k] + first;
	}
}

// Forces of the current state with the workgroup engine's lanes per
// nucleotide (after uploads, list rebuilds and for observables), so that
// runs continue from exactly the forces the engine itself computes.
template<bool SPLIT, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_persist_forces(PersistArgs a, double *esplit, double *stress) {
	const oxh_sys &S = a.systems[blockIdx.x];
	const int K = persist_lanes(S.n, BLOCK);
	const int tid = threadIdx.x;
	if(tid < S.n * K)
		particle_forces<SPLIT, true>(K, S.first + tid / K, tid % K, a.N, a.pos, a.q64, a.q32, a.bpos, a.bonds, a.pinfo, a.nbr, a.nnbr,
				a.max_neigh, a.systems, a.state, a.F, a.T, esplit, a.hbcount, stress);
}

PersistArgs make_args(const Batch &b, int nsteps) {
	PersistArgs a;
	a.N = b.N;
	a.max_neigh = b.max_neigh;
	a.nsteps = nsteps;
	a.cell_cap = b.cell_cap;
	a.pos = b.pos;
	a.vel = b.vel;
	a.Lm = b.Lm;
	a.q64 = b.q64;
	a.rlist = b.rlist;
	a.q32 = b.q32;
	a.F = b.F;
	a.T = b.T;
	a.bpos = b.bpos;
	a.bonds = b.bonds;
	a.pinfo = b.pinfo;
	a.orig = b.orig;
	a.pstep = b.pstep;
	a.nbr = b.nbr;
	a.nnbr = b.nnbr;
	a.systems = b.d_sys;
	a.state = b.d_state;
	a.cell_count = b.cell_count;
	a.cell_items = b.cell_items;
	a.pos32 = b.pos32;
	a.hbcount = b.want_hbcount ? b.hbcount : nullptr;
	a.limit = b.limit;
	return a;
}

}  // namespace

size_t persist_lds_bytes(int nmax) {
	const size_t per = 5 * sizeof(double4) + 4 * sizeof(float4) + sizeof(long long) + 3 * sizeof(int2) + sizeof(int);
	return (size_t) nmax * per + (size_t) nmax * (nmax > 1 ? nmax - 1 : 1) * sizeof(int);
}

void launch_persist(hipStream_t st, const Batch &b, int nsteps, int block) {
	if(b.N == 0 || nsteps <= 0) return;
	if(b.persist_lds_n > 0) {
		const size_t bytes = persist_lds_bytes(b.persist_lds_n);
		if(block == 256) k_persist_lds<256><<<b.nsys, 256, bytes, st>>>(make_args(b, nsteps), b.persist_lds_n);
		else k_persist_lds<512><<<b.nsys, 512, bytes, st>>>(make_args(b, nsteps), b.persist_lds_n);
		return;
	}
	if(block == 256) k_persist<256><<<b.nsys, 256, 0, st>>>(make_args(b, nsteps));
	else k_persist<512><<<b.nsys, 512, 0, st>>>(make_args(b, nsteps));
}

void launch_persist_forces(hipStream_t st, const Batch &b, bool split, int block) {
	if(b.N == 0) return;
	PersistArgs a = make_args(b, 0);
	if(block == 256) {
		if(split) k_persist_forces<true, 256><<<b.nsys, 256, 0, st>>>(a, b.esplit, b.stress);
		else k_persist_forces<false, 256><<<b.nsys, 256, 0, st>>>(a, nullptr, nullptr);
	}
	else {
		if(split) k_persist_forces<true, 512><<<b.nsys, 512, 0, st>>>(a, b.esplit, b.stress);
		else k_persist_forces<false, 512><<<b.nsys, 512, 0, st>>>(a, nullptr, nullptr);
	}
}

}  // namespace oxh
