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

namespace oxh {

enum { RNG_LANGEVIN = 0, RNG_BROWNIAN = 2, RNG_REFRESH = 4, RNG_BUSSI = 6 };

__device__ __forceinline__ uint4 rng(const oxh_sys &S, uint32_t idx, long long step, uint32_t stream) {
	return philox4x32_10(make_uint4(idx, (uint32_t) step, (uint32_t) ((unsigned long long) step >> 32), stream),
			make_uint2(S.seed_lo, S.seed_hi));
}

__device__ __forceinline__ void normals4(const oxh_sys &S, uint32_t idx, long long step, uint32_t stream, double n[4]) {
	uint4 w = rng(S, idx, step, stream);
	box_muller(w.x, w.y, n[0], n[1]);
	box_muller(w.z, w.w, n[2], n[3]);
}

// quaternion (x, y, z, w) Hamilton product a * b
__device__ __forceinline__ double4 qmul(double4 a, double4 b) {
	return make_double4(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
			a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}

// FP32 mirror of the orientation, and the backbone-site position wrapped
// into the box in FP64 and rounded to FP32, with the particle's
// Debye-Hueckel charge factor in .w (0.5 at strand ends when
// dh_half_charged_ends, src/Interactions/DNA2Interaction.cpp:162-165)
__device__ __forceinline__ void write_mirror(int i, double4 r, double4 q, const oxh_sys &S, int2 bd, float4 *q32, float4 *bpos) {
	q32[i] = make_float4((float) q.x, (float) q.y, (float) q.z, (float) q.w);
	double sqx = q.x * q.x, sqy = q.y * q.y, sqz = q.z * q.z, sqw = q.w * q.w;
	double xy = q.x * q.y, xz = q.x * q.z, xw = q.x * q.w, yz = q.y * q.z, yw = q.y * q.w, zw = q.z * q.w;
	double a1x = sqx - sqy - sqz + sqw, a1y = 2. * (xy + zw), a1z = 2. * (xz - yw);
	double a2x = 2. * (xy - zw), a2y = -sqx + sqy - sqz + sqw, a2z = 2. * (yz + xw);
	const double b1 = (double) POS_MM_BACK1, b2 = (double) POS_MM_BACK2;
	float cf = (S.dh_half && (bd.x < 0 || bd.y < 0)) ? 0.5f : 1.f;
	double bx = r.x + b1 * a1x + b2 * a2x, by = r.y + b1 * a1y + b2 * a2y, bz = r.z + b1 * a1z + b2 * a2z;
	bx -= S.box[0] * floor(bx * S.inv_box[0]);
	by -= S.box[1] * floor(by * S.inv_box[1]);
	bz -= S.box[2] * floor(bz * S.inv_box[2]);
	bpos[i] = make_float4((float) bx, (float) by, (float) bz, cf);
}

__global__ void k_prepare_mirror(int N, const double4 *__restrict__ pos, const double4 *__restrict__ q64, const int2 *__restrict__ bonds,
		const int2 *__restrict__ pinfo, const oxh_sys *__restrict__ systems, float4 *__restrict__ q32, float4 *__restrict__ bpos) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= N) return;
	write_mirror(i, pos[i], q64[i], systems[pinfo[i].y], bonds[i], q32, bpos);
}

void launch_prepare_mirror(hipStream_t st, const Batch &b) {
	if(b.N == 0) return;
	k_prepare_mirror<<<(b.N + 255) / 256, 256, 0, st>>>(b.N, b.pos, b.q64, b.bonds, b.pinfo, b.d_sys, b.q32, b.bpos);
}

__global__ void __launch_bounds__(256) k_first_step(int N, double4 *__restrict__ pos, double4 *__restrict__ vel, double4 *__restrict__ Lm,
		double4 *__restrict__ q64, float4 *__restrict__ q32, float4 *__restrict__ bpos, const float4 *__restrict__ F,
		const float4 *__restrict__ T, const double4 *__restrict__ rlist, const int2 *__restrict__ bonds, const int2 *__restrict__ pinfo,
		long long *__restrict__ pstep, const oxh_sys *__restrict__ systems, oxh_sys_state *__restrict__ st, int ncell,
		int *__restrict__ cell_count) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	// cell counters start every step at zero; they are only filled (after
	// this kernel) in steps that rebuild the lists
	if(i < ncell) cell_count[i] = 0;
	if(i >= N) return;
	const int s = pinfo[i].y;
	const oxh_sys &S = systems[s];
	const double hdt = 0.5 * S.dt;

	float4 f = F[i];
	double4 v = vel[i];
	v.x += (double) f.x * hdt;
	v.y += (double) f.y * hdt;
	v.z += (double) f.z * hdt;
	vel[i] = v;
	double4 r = pos[i];
	r.x += v.x * S.dt;
	r.y += v.y * S.dt;
	r.z += v.z * S.dt;
	pos[i] = r;

	float4 t = T[i];
	double4 L = Lm[i];
	L.x += (double) t.x * hdt;
	L.y += (double) t.y * hdt;
	L.z += (double) t.z * hdt;
	Lm[i] = L;
	double4 q = q64[i];
	double norm = sqrt(L.x * L.x + L.y * L.y + L.z * L.z);
	if(norm > 0.) {
		double sh, ch;
		sincos(0.5 * S.dt * norm, &sh, &ch);
		double k = sh / norm;
		q = qmul(q, make_double4(L.x * k, L.y * k, L.z * k, ch));
		double qn = rsqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
		q.x *= qn;
		q.y *= qn;
		q.z *= qn;
		q.w *= qn;
		q64[i] = q;
	}
	write_mirror(i, r, q, S, bonds[i], q32, bpos);
	pstep[i] += 1;

	double4 rl = rlist[i];
	double dx = r.x - rl.x, dy = r.y - rl.y, dz = r.z - rl.z;
	if(dx * dx + dy * dy + dz * dz > S.skin2) st[s].need_rebuild = 1;
}

void launch_first_step(hipStream_t st, const Batch &b) {
	int n = std::max(b.N, b.ncell_total);
	k_first_step<<<(n + 255) / 256, 256, 0, st>>>(b.N, b.pos, b.vel, b.Lm, b.q64, b.q32, b.bpos, b.F, b.T, b.rlist, b.bonds, b.pinfo,
			b.pstep, b.d_sys, b.d_state, b.ncell_total, b.cell_count);
}

// second half kick + Langevin / Brownian thermostats; the first particle
// of each system closes the system's list-rebuild flag for this step
__global__ void __launch_bounds__(256) k_second_step(int N, double4 *__restrict__ vel, double4 *__restrict__ Lm,
		const float4 *__restrict__ F, const float4 *__restrict__ T, const int2 *__restrict__ pinfo, const int *__restrict__ orig,
		const long long *__restrict__ pstep, const oxh_sys *__restrict__ systems, oxh_sys_state *__restrict__ st) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= N) return;
	const int s = pinfo[i].y;
	const oxh_sys &S = systems[s];
	const double hdt = 0.5 * S.dt;
	float4 f = F[i], t = T[i];
	double4 v = vel[i], L = Lm[i];
	v.x += (double) f.x * hdt;
	v.y += (double) f.y * hdt;
	v.z += (double) f.z * hdt;
	L.x += (double) t.x * hdt;
	L.y += (double) t.y * hdt;
	L.z += (double) t.z * hdt;

	const long long step = pstep[i] - 1;    // index of the step being completed
	const uint32_t id = (uint32_t) orig[i];
	if(S.thermostat == 1) {
		// src/Backends/Thermostats/LangevinThermostat.cpp:78-85
		double n[8];
		normals4(S, id, step, RNG_LANGEVIN, n);
		normals4(S, id, step, RNG_LANGEVIN + 1, n + 4);
		v.x += S.dt * (-S.gamma_t * v.x + n[0] * S.resc_t);
		v.y += S.dt * (-S.gamma_t * v.y + n[1] * S.resc_t);
		v.z += S.dt * (-S.gamma_t * v.z + n[2] * S.resc_t);
		L.x += S.dt * (-S.gamma_r * L.x + n[3] * S.resc_r);
		L.y += S.dt * (-S.gamma_r * L.y + n[4] * S.resc_r);
		L.z += S.dt * (-S.gamma_r * L.z + n[5] * S.resc_r);
	}
	else if(S.thermostat == 3 && step % S.newtonian_steps == 0) {
		// src/Backends/Thermostats/BrownianThermostat.cpp:58-68
		uint4 u = rng(S, id, step, RNG_BROWNIAN);
		double n[4];
		if(u01(u.x) < S.pt) {
			normals4(S, id, step, RNG_BROWNIAN + 1, n);
			v.x = n[0] * S.resc_brown;
			v.y = n[1] * S.resc_brown;
			v.z = n[2] * S.resc_brown;
		}
		if(u01(u.y) < S.pr) {
			normals4(S, id, step, RNG_BROWNIAN + 2 + 4, n);
			L.x = n[0] * S.resc_brown;
			L.y = n[1] * S.resc_brown;
			L.z = n[2] * S.resc_brown;
		}
	}
	vel[i] = v;
	Lm[i] = L;
	if(i == S.first && st[s].need_rebuild) {
		st[s].need_rebuild = 0;
		st[s].n_rebuilds += 1;
	}
}

void launch_second_step(hipStream_t st, const Batch &b) {
	k_second_step<<<(b.N + 255) / 256, 256, 0, st>>>(b.N, b.vel, b.Lm, b.F, b.T, b.pinfo, b.orig, b.pstep, b.d_sys, b.d_state);
}

void launch_end_step(hipStream_t, const Batch &) {}

// -------------------------------------------------------- FP64 reductions
// per system: U = sum F.w, U_hb = sum T.w, K_t = 1/2 sum v^2, K_r = 1/2 sum L^2,
// sum v (3 components), N.  Blocks never straddle systems; the final sum
// runs over a system's blocks in order, so results are reproducible.
__global__ void __launch_bounds__(256) k_reduce_partial(const int3 *__restrict__ blocks, const float4 *__restrict__ F,
		const float4 *__restrict__ T, const double4 *__restrict__ vel, const double4 *__restrict__ Lm, double *__restrict__ partial) {
	__shared__ double sh[RED_W][256];
	int3 bl = blocks[blockIdx.x];
	double acc[RED_W] = {0, 0, 0, 0, 0, 0, 0, 0};
	for(int i = bl.y + threadIdx.x; i < bl.z; i += blockDim.x) {
		float4 f = F[i], t = T[i];
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
	k_reduce_partial<<<b.nred, 256, 0, st>>>(b.red_blocks, b.F, b.T, b.vel, b.Lm, b.red_partial);
	k_reduce_final<<<(b.nsys + 63) / 64, 64, 0, st>>>(b.nsys, b.red_sys_blocks, b.red_partial, b.red_out);
}

// ------------------------------------------------------------------ Bussi
// Gamma(a, 1) variate, a >= 1, after G. Marsaglia and W. W. Tsang,
// "A simple method for generating gamma variables", ACM TOMS 26 (2000).
__device__ double gamma_mt(const oxh_sys &S, long long step, uint32_t stream, double a) {
	double d = a - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
	for(uint32_t it = 0; it < 1000; it++) {
		uint4 w = rng(S, 0xFFFF0000u + it, step, stream);
		double x, y;
		box_muller(w.x, w.y, x, y);
		double v = 1.0 + c * x;
		if(v <= 0.) continue;
		v = v * v * v;
		double u = u01(w.z);
		if(log(u) < 0.5 * x * x + d - d * v + d * log(v)) return d * v;
	}
	return d;
}

// sum of n squared standard normals (chi-squared with n degrees of freedom)
__device__ double sum_noises(const oxh_sys &S, long long step, uint32_t stream, int n) {
	if(n <= 0) return 0.;
	if(n == 1) {
		double a, b;
		uint4 w = rng(S, 0xFFFE0000u, step, stream);
		box_muller(w.x, w.y, a, b);
		return a * a;
	}
	if(n % 2 == 0) return 2.0 * gamma_mt(S, step, stream, n / 2);
	double a, b;
	uint4 w = rng(S, 0xFFFE0000u, step, stream);
	box_muller(w.x, w.y, a, b);
	return 2.0 * gamma_mt(S, step, stream, (n - 1) / 2) + a * a;
}

// src/Backends/Thermostats/BussiThermostat.cpp:45-52
__device__ double bussi_K(const oxh_sys &S, long long step, uint32_t stream, double K, int dof) {
	double K_target = 0.5 * dof * S.T;
	double a, b;
	uint4 w = rng(S, 0xFFFD0000u, step, stream);
	box_muller(w.x, w.y, a, b);
	double rr = a;
	double c = S.bussi_c;
	double dK = (1.0 - c) * (K_target * (sum_noises(S, step, stream + 1, dof - 1) + rr * rr) / dof - K) +
			2.0 * rr * sqrt(K * K_target / dof * (1.0 - c) * c);
	return K + dK;
}

// One pass over velocities and angular momenta: every block reduces its
// chunk (fixed tree) into partial[]; the last block of a system to finish
// (device-scope counter) sums the system's partials in block order and
// advances the thermostat's kinetic energies
// (src/Backends/Thermostats/BussiThermostat.cpp:54-90).
__global__ void __launch_bounds__(256) k_bussi_reduce(const int3 *__restrict__ blocks, const int2 *__restrict__ sys_blocks,
		const double4 *__restrict__ vel, const double4 *__restrict__ Lm, const oxh_sys *__restrict__ systems,
		oxh_sys_state *__restrict__ st, const long long *__restrict__ pstep, double *__restrict__ partial, int *__restrict__ done,
		double4 *__restrict__ factors, int nsys) {
	__shared__ double sh[5][256];
	__shared__ int last;
	const int3 bl = blocks[blockIdx.x];
	const int s = bl.x;
	const oxh_sys &S = systems[s];
	if(S.thermostat != 2) return;
	const long long step = pstep[S.first] - 1;
	if(step % S.newtonian_steps != 0) {
		if(threadIdx.x == 0 && blockIdx.x == sys_blocks[s].x) factors[s] = make_double4(1.0, 1.0, 0.0, 0.0);
		return;
	}
	double acc[5] = {0, 0, 0, 0, 0};
	for(int i = bl.y + threadIdx.x; i < bl.z; i += blockDim.x) {
		double4 v = vel[i], L = Lm[i];
		acc[0] += 0.5 * (v.x * v.x + v.y * v.y + v.z * v.z);
		acc[1] += 0.5 * (L.x * L.x + L.y * L.y + L.z * L.z);
		acc[2] += v.x;
		acc[3] += v.y;
		acc[4] += v.z;
	}
	for(int k = 0; k < 5; k++) sh[k][threadIdx.x] = acc[k];
	__syncthreads();
	for(int off = 128; off > 0; off >>= 1) {
		if(threadIdx.x < off)
			for(int k = 0; k < 5; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + off];
		__syncthreads();
	}
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
	const double n = S.n;
	const double vcx = r[2] / n, vcy = r[3] / n, vcz = r[4] / n;
	const double K_now_t = r[0] - 0.5 * n * (vcx * vcx + vcy * vcy + vcz * vcz);
	const double K_now_r = r[1];
	const int dof_t = 3 * (S.n - 1), dof_r = 3 * S.n;
	st[s].K_t = bussi_K(S, step, RNG_BUSSI, st[s].K_t, dof_t);
	st[s].K_r = bussi_K(S, step, RNG_BUSSI + 4, st[s].K_r, dof_r);
	const double ft = K_now_t > 0. ? sqrt(st[s].K_t / K_now_t) : 1.0;
	const double fr = K_now_r > 0. ? sqrt(st[s].K_r / K_now_r) : 1.0;
	factors[s] = make_double4(ft, fr, 1.0, 0.0);
	factors[nsys + s] = make_double4(vcx, vcy, vcz, 0.0);
}

__global__ void k_bussi_rescale(int N, double4 *__restrict__ vel, double4 *__restrict__ Lm, const int2 *__restrict__ pinfo,
		const double4 *__restrict__ factors, int nsys, const oxh_sys *__restrict__ systems) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= N) return;
	int s = pinfo[i].y;
	if(systems[s].thermostat != 2) return;
	double4 f = factors[s];
	if(f.z == 0.0) return;
	double4 vc = factors[nsys + s];
	double4 v = vel[i];
	v.x = (v.x - vc.x) * f.x + vc.x;
	v.y = (v.y - vc.y) * f.x + vc.y;
	v.z = (v.z - vc.z) * f.x + vc.z;
	vel[i] = v;
	double4 L = Lm[i];
	L.x *= f.y;
	L.y *= f.y;
	L.z *= f.y;
	Lm[i] = L;
}

void launch_bussi(hipStream_t st, const Batch &b) {
	double4 *factors = (double4 *) b.scratch;
	k_bussi_reduce<<<b.nred, 256, 0, st>>>(b.red_blocks, b.red_sys_blocks, b.vel, b.Lm, b.d_sys, b.d_state, b.pstep, b.red_partial,
			b.red_done, factors, b.nsys);
	k_bussi_rescale<<<(b.N + 255) / 256, 256, 0, st>>>(b.N, b.vel, b.Lm, b.pinfo, factors, b.nsys, b.d_sys);
}

// initial velocities ~ N(0, T) per component (src/Backends/MDBackend.cpp:140-165)
__global__ void k_refresh_vel(int N, double4 *__restrict__ vel, double4 *__restrict__ Lm, const int2 *__restrict__ pinfo,
		const int *__restrict__ orig, const oxh_sys *__restrict__ systems) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= N) return;
	const oxh_sys &S = systems[pinfo[i].y];
	double n[8];
	normals4(S, (uint32_t) orig[i], -1, RNG_REFRESH, n);
	normals4(S, (uint32_t) orig[i], -1, RNG_REFRESH + 1, n + 4);
	double sT = sqrt(S.T);
	vel[i] = make_double4(n[0] * sT, n[1] * sT, n[2] * sT, 0.0);
	Lm[i] = make_double4(n[3] * sT, n[4] * sT, n[5] * sT, 0.0);
}

void launch_refresh_vel(hipStream_t st, const Batch &b) {
	if(b.N == 0) return;
	k_refresh_vel<<<(b.N + 255) / 256, 256, 0, st>>>(b.N, b.vel, b.Lm, b.pinfo, b.orig, b.d_sys);
}

}  // namespace oxh
