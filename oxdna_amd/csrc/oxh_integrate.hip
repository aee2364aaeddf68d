// Velocity-Verlet integration, thermostats and reductions on gfx950.
//
// State is FP64 (positions, velocities, angular momenta, orientation
// quaternions); forces and torques arrive in FP32 and are accumulated into
// the FP64 state.  Unit mass and moment of inertia, as the reference
// (src/Backends/MD_CPUBackend.cpp:67-180, src/CUDA/Backends/CUDA_mixed.cuh:
// 8-93).  The orientation update right-multiplies the body quaternion by the
// rotation of angle dt*|L| about L (the quaternion form of the reference
// CPU's Rodrigues matrix, src/Backends/MD_CPUBackend.cpp:107-131).
//
// Thermostats follow src/Backends/Thermostats/{Langevin,Bussi,Brownian}
// Thermostat.cpp; their noise comes from Philox4x32-10 keyed by the
// system's seed and counted by (particle, step, stream), so a system's
// trajectory does not depend on which other systems share the batch or on
// the particle order.  Bussi's kinetic-energy reductions are FP64.
#include <algorithm>

#include "oxh_device.h"
#include "oxh_internal.h"
#include "oxh_step.h"

namespace oxh {

__global__ void k_prepare_mirror(int N, const double4 *__restrict__ pos, const double4 *__restrict__ q64, const int2 *__restrict__ bonds,
		const int2 *__restrict__ pinfo, const oxh_sys *__restrict__ systems, float4 *__restrict__ q32, float4 *__restrict__ bpos,
		double4 *__restrict__ back64) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= N) return;
	const oxh_sys &S = systems[pinfo[i].y];
	if(padding(i, S)) return;
	write_mirror(i, pos[i], q64[i], S, bonds[i], q32, bpos, back64);
}

void launch_prepare_mirror(hipStream_t st, const Batch &b) {
	if(b.N == 0) return;
	k_prepare_mirror<<<(b.N + 255) / 256, 256, 0, st>>>(b.N, b.pos, b.q64, b.bonds, b.pinfo, b.d_sys, b.q32, b.bpos, b.back64);
}

template<class FT>
__global__ void __launch_bounds__(256) k_first_step(int N, double4 *__restrict__ pos, double4 *__restrict__ vel, double4 *__restrict__ Lm,
		double4 *__restrict__ q64, float4 *__restrict__ q32, float4 *__restrict__ bpos, double4 *__restrict__ back64,
		const FT *__restrict__ F, const FT *__restrict__ T, const double4 *__restrict__ rlist, const int2 *__restrict__ bonds, const int2 *__restrict__ pinfo,
		long long *__restrict__ pstep, const oxh_sys *__restrict__ systems, oxh_sys_state *__restrict__ st, int ncell,
		int *__restrict__ cell_count, int k, const int *__restrict__ limit, int *__restrict__ any, const float4 *__restrict__ blist) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	// cell counters start every step at zero; they are only filled (after
	// this kernel) in steps that rebuild the lists
	if(i < ncell) cell_count[i] = 0;
	if(i >= N) return;
	const int s = pinfo[i].y;
	const oxh_sys &S = systems[s];
	if(idle(i, S, s, k, limit)) return;
	if(first_step_particle(i, S, pos, vel, Lm, q64, q32, bpos, back64, F, T, rlist, bonds, pstep, blist)) {
		st[s].need_rebuild = 1;
		*any = 1;
	}
}

void launch_first_step(hipStream_t st, const Batch &b, int k) {
	int n = std::max(b.N, b.ncell_total);
	if(b.fp64)
		k_first_step<<<(n + 255) / 256, 256, 0, st>>>(b.N, b.pos, b.vel, b.Lm, b.q64, b.q32, b.bpos, b.back64, b.F64, b.T64, b.rlist, b.bonds, b.pinfo,
				b.pstep, b.d_sys, b.d_state, b.ncell_total, b.cell_count, k, b.limit, b.any_rebuild, b.blist);
	else
		k_first_step<<<(n + 255) / 256, 256, 0, st>>>(b.N, b.pos, b.vel, b.Lm, b.q64, b.q32, b.bpos, b.back64, b.F, b.T, b.rlist, b.bonds, b.pinfo,
				b.pstep, b.d_sys, b.d_state, b.ncell_total, b.cell_count, k, b.limit, b.any_rebuild, b.blist);
}

// second half kick + Langevin / Brownian thermostats; the first particle
// of each system closes the system's list-rebuild flag for this step
template<class FT>
__global__ void __launch_bounds__(256) k_second_step(int N, double4 *__restrict__ vel, double4 *__restrict__ Lm,
		const FT *__restrict__ F, const FT *__restrict__ T, const int2 *__restrict__ pinfo, const int *__restrict__ orig,
		const long long *__restrict__ pstep, const oxh_sys *__restrict__ systems, oxh_sys_state *__restrict__ st, int k,
		const int *__restrict__ limit) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= N) return;
	const int s = pinfo[i].y;
	const oxh_sys &S = systems[s];
	if(idle(i, S, s, k, limit)) return;
	second_step_particle(i, S, vel, Lm, F, T, orig, pstep);
}

void launch_second_step(hipStream_t st, const Batch &b, int k) {
	if(b.fp64)
		k_second_step<<<(b.N + 255) / 256, 256, 0, st>>>(b.N, b.vel, b.Lm, b.F64, b.T64, b.pinfo, b.orig, b.pstep, b.d_sys, b.d_state, k,
				b.limit);
	else
		k_second_step<<<(b.N + 255) / 256, 256, 0, st>>>(b.N, b.vel, b.Lm, b.F, b.T, b.pinfo, b.orig, b.pstep, b.d_sys, b.d_state, k,
				b.limit);
}

// The second half of step k and the first half of step k + 1 in one pass
// over the particles (when no Bussi activation falls in between): the same
// per-particle functions in the same order as k_second_step followed by
// k_first_step.
template<class FT>
__global__ void __launch_bounds__(256) k_second_first(int N, double4 *__restrict__ pos, double4 *__restrict__ vel,
		double4 *__restrict__ Lm, double4 *__restrict__ q64, float4 *__restrict__ q32, float4 *__restrict__ bpos,
		double4 *__restrict__ back64, const FT *__restrict__ F, const FT *__restrict__ T, const double4 *__restrict__ rlist,
		const int2 *__restrict__ bonds, const int2 *__restrict__ pinfo, const int *__restrict__ orig, long long *__restrict__ pstep, const oxh_sys *__restrict__ systems,
		oxh_sys_state *__restrict__ st, int ncell, int *__restrict__ cell_count, int k, const int *__restrict__ limit,
		int *__restrict__ any, const float4 *__restrict__ blist) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i < ncell) cell_count[i] = 0;
	if(i >= N) return;
	const int s = pinfo[i].y;
	const oxh_sys &S = systems[s];
	if(padding(i, S)) return;
	if(!(limit && k >= limit[s])) second_step_particle(i, S, vel, Lm, F, T, orig, pstep);
	if(!(limit && k + 1 >= limit[s])) {
		if(first_step_particle(i, S, pos, vel, Lm, q64, q32, bpos, back64, F, T, rlist, bonds, pstep, blist)) {
			st[s].need_rebuild = 1;
			*any = 1;
		}
	}
}

void launch_second_first(hipStream_t st, const Batch &b, int k) {
	int n = std::max(b.N, b.ncell_total);
	if(b.fp64)
		k_second_first<<<(n + 255) / 256, 256, 0, st>>>(b.N, b.pos, b.vel, b.Lm, b.q64, b.q32, b.bpos, b.back64, b.F64, b.T64, b.rlist, b.bonds,
				b.pinfo, b.orig, b.pstep, b.d_sys, b.d_state, b.ncell_total, b.cell_count, k, b.limit, b.any_rebuild, b.blist);
	else
		k_second_first<<<(n + 255) / 256, 256, 0, st>>>(b.N, b.pos, b.vel, b.Lm, b.q64, b.q32, b.bpos, b.back64, b.F, b.T, b.rlist, b.bonds, b.pinfo,
				b.orig, b.pstep, b.d_sys, b.d_state, b.ncell_total, b.cell_count, k, b.limit, b.any_rebuild, b.blist);
}

// -------------------------------------------------------- FP64 reductions
// per system: U = sum F.w, U_hb = sum T.w, K_t = 1/2 sum v^2, K_r = 1/2 sum L^2,
// sum v (3 components), N.  Blocks never straddle systems; the final sum
// runs over a system's blocks in order, so results are reproducible.
template<class FT>
__global__ void __launch_bounds__(256) k_reduce_partial(const int3 *__restrict__ blocks, const FT *__restrict__ F,
		const FT *__restrict__ T, const double4 *__restrict__ vel, const double4 *__restrict__ Lm, double *__restrict__ partial) {
	__shared__ double sh[RED_W][256];
	int3 bl = blocks[blockIdx.x];
	if(bl.x < 0) bl.y = bl.z = 0;   // unused block of a slot pool
	double acc[RED_W] = {0, 0, 0, 0, 0, 0, 0, 0};
	for(int i = bl.y + threadIdx.x; i < bl.z; i += blockDim.x) {
		const FT f = F[i], t = T[i];
		double4 v = vel[i], L = Lm[i];
		acc[0] += (double) f.w;
		acc[1] += (double) t.w;
		acc[2] += 0.5 * (v.x * v.x + v.y * v.y + v.z * v.z);
		acc[3] += 0.5 * (L.x * L.x + L.y * L.y + L.z * L.z);
		acc[4] += v.x;
		acc[5] += v.y;
		acc[6] += v.z;
		acc[7] += 1.0;
	}
	for(int k = 0; k < RED_W; k++) sh[k][threadIdx.x] = acc[k];
	__syncthreads();
	for(int off = 128; off > 0; off >>= 1) {
		if(threadIdx.x < off)
			for(int k = 0; k < RED_W; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + off];
		__syncthreads();
	}
	if(threadIdx.x < RED_W) partial[(size_t) blockIdx.x * RED_W + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_reduce_final(int nsys, const int2 *__restrict__ sys_blocks, const double *__restrict__ partial, double *__restrict__ out) {
	int s = blockIdx.x * blockDim.x + threadIdx.x;
	if(s >= nsys) return;
	int2 sb = sys_blocks[s];
	for(int k = 0; k < RED_W; k++) {
		double acc = 0.;
		for(int b = sb.x; b < sb.x + sb.y; b++) acc += partial[(size_t) b * RED_W + k];
		out[(size_t) s * RED_W + k] = acc;
	}
}

void launch_reduce(hipStream_t st, const Batch &b) {
	if(b.nred == 0) return;
	if(b.fp64) k_reduce_partial<<<b.nred, 256, 0, st>>>(b.red_blocks, b.F64, b.T64, b.vel, b.Lm, b.red_partial);
	else k_reduce_partial<<<b.nred, 256, 0, st>>>(b.red_blocks, b.F, b.T, b.vel, b.Lm, b.red_partial);
	k_reduce_final<<<(b.nsys + 63) / 64, 64, 0, st>>>(b.nsys, b.red_sys_blocks, b.red_partial, b.red_out);
}

// One pass over velocities and angular momenta: every block reduces its
// chunk (fixed tree) into partial[]; the last block of a system to finish
// (device-scope counter) sums the system's partials in block order and
// advances the thermostat's kinetic energies
// (src/Backends/Thermostats/BussiThermostat.cpp:54-90).
__global__ void __launch_bounds__(256) k_bussi_reduce(const int3 *__restrict__ blocks, const int2 *__restrict__ sys_blocks,
		const double4 *__restrict__ vel, const double4 *__restrict__ Lm, const oxh_sys *__restrict__ systems,
		oxh_sys_state *__restrict__ st, const long long *__restrict__ pstep, double *__restrict__ partial, int *__restrict__ done,
		double4 *__restrict__ factors, int nsys, int k, const int *__restrict__ limit) {
	__shared__ double sh[5][256];
	__shared__ int last;
	const int3 bl = blocks[blockIdx.x];
	const int s = bl.x;
	if(s < 0) return;   // unused block of a slot pool
	const oxh_sys &S = systems[s];
	if(S.thermostat != 2) return;
	const long long step = pstep[S.first] - 1;
	if(step % S.newtonian_steps != 0 || (limit && k >= limit[s])) {
		if(threadIdx.x == 0 && blockIdx.x == sys_blocks[s].x) factors[s] = make_double4(1.0, 1.0, 0.0, 0.0);
		return;
	}
	bussi_chunk_sums(bl.y, bl.z, vel, Lm, sh);
	if(threadIdx.x == 0) {
		for(int k = 0; k < 5; k++) partial[(size_t) blockIdx.x * RED_W + k] = sh[k][0];
		__threadfence();
		const int2 sb = sys_blocks[s];
		last = (atomicAdd(&done[s], 1) == sb.y - 1);
	}
	__syncthreads();
	if(!last || threadIdx.x != 0) return;
	__threadfence();
	const int2 sb = sys_blocks[s];
	double r[5] = {0, 0, 0, 0, 0};
	for(int b = sb.x; b < sb.x + sb.y; b++)
		for(int k = 0; k < 5; k++) r[k] += __builtin_nontemporal_load(&partial[(size_t) b * RED_W + k]);
	done[s] = 0;
	double ft, fr, vc[3];
	bussi_update(S, st[s], step, r, ft, fr, vc);
	factors[s] = make_double4(ft, fr, 1.0, 0.0);
	factors[nsys + s] = make_double4(vc[0], vc[1], vc[2], 0.0);
}

__global__ void k_bussi_rescale(int N, double4 *__restrict__ vel, double4 *__restrict__ Lm, const int2 *__restrict__ pinfo,
		const double4 *__restrict__ factors, int nsys, const oxh_sys *__restrict__ systems) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= N) return;
	int s = pinfo[i].y;
	if(systems[s].thermostat != 2 || padding(i, systems[s])) return;
	double4 f = factors[s];
	if(f.z == 0.0) return;
	double4 vc = factors[nsys + s];
	bussi_rescale_particle(i, vel, Lm, f.x, f.y, vc.x, vc.y, vc.z);
}

void launch_bussi(hipStream_t st, const Batch &b, int k) {
	double4 *factors = (double4 *) b.scratch;
	k_bussi_reduce<<<b.nred, 256, 0, st>>>(b.red_blocks, b.red_sys_blocks, b.vel, b.Lm, b.d_sys, b.d_state, b.pstep, b.red_partial,
			b.red_done, factors, b.nsys, k, b.limit);
	k_bussi_rescale<<<(b.N + 255) / 256, 256, 0, st>>>(b.N, b.vel, b.Lm, b.pinfo, factors, b.nsys, b.d_sys);
}

// initial velocities ~ N(0, T) per component (src/Backends/MDBackend.cpp:140-165)
__global__ void k_refresh_vel(int i0, int n_, double4 *__restrict__ vel, double4 *__restrict__ Lm, const int2 *__restrict__ pinfo,
		const int *__restrict__ orig, const oxh_sys *__restrict__ systems) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n_) return;
	i += i0;
	const oxh_sys &S = systems[pinfo[i].y];
	if(padding(i, S)) return;
	double n[8];
	normals4(S, (uint32_t) orig[i], -1, RNG_REFRESH, n);
	normals4(S, (uint32_t) orig[i], -1, RNG_REFRESH + 1, n + 4);
	double sT = sqrt(S.T);
	vel[i] = make_double4(n[0] * sT, n[1] * sT, n[2] * sT, 0.0);
	Lm[i] = make_double4(n[3] * sT, n[4] * sT, n[5] * sT, 0.0);
}

void launch_refresh_vel(hipStream_t st, const Batch &b, int i0, int n) {
	if(n <= 0) return;
	k_refresh_vel<<<(n + 255) / 256, 256, 0, st>>>(i0, n, b.vel, b.Lm, b.pinfo, b.orig, b.d_sys);
}

}  // namespace oxh
