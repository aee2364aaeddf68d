// Per-particle integration steps and thermostat helpers on gfx950, shared
// by the one-kernel-per-phase path (oxh_integrate.hip) and the
// workgroup-resident engine (oxh_persist.hip).
#pragma once

#include "oxh_device.h"

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
// back64 (all-FP64 mode, else nullptr): the FP64 backbone offset
// POS_MM_BACK1 a1 + POS_MM_BACK2 a2 of the particle, for the FP64 far pairs
__device__ __forceinline__ float4 write_mirror(int i, double4 r, double4 q, const oxh_sys &S, int2 bd, float4 *q32, float4 *bpos,
		double4 *back64) {
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
	const float4 bp = make_float4((float) bx, (float) by, (float) bz, cf);
	bpos[i] = bp;
	if(back64) back64[i] = make_double4(b1 * a1x + b2 * a2x, b1 * a1y + b2 * a2y, b1 * a1z + b2 * a2z, 0.0);
	return bp;
}

// With far pairs kept by their backbone sites: true when the wrapped FP32
// backbone site b has moved more than sqrt(sback2f) from blist, its place
// at the last list build (nearest image, so wrapping does not count)
__device__ __forceinline__ bool back_moved(float4 b, float4 bl, const oxh_sys &S) {
	float dx = b.x - bl.x, dy = b.y - bl.y, dz = b.z - bl.z;
	dx -= S.boxf[0] * rintf(dx * S.inv_boxf[0]);
	dy -= S.boxf[1] * rintf(dy * S.inv_boxf[1]);
	dz -= S.boxf[2] * rintf(dz * S.inv_boxf[2]);
	return dx * dx + dy * dy + dz * dz > S.sback2f;
}

// first half kick, drift and rotation of particle i, FP32 mirror, step
// counter; returns true when the particle moved more than the skin since
// the last list build, or its backbone site more than its bound when far
// pairs are kept by backbone sites (blist: the sites at the build; may be
// null) (src/CUDA/Backends/CUDA_mixed.cuh:23-69)
template<class FT>
__device__ __forceinline__ bool first_step_particle(int i, const oxh_sys &S, double4 *__restrict__ pos, double4 *__restrict__ vel,
		double4 *__restrict__ Lm, double4 *__restrict__ q64, float4 *__restrict__ q32, float4 *__restrict__ bpos,
		double4 *__restrict__ back64, const FT *__restrict__ F, const FT *__restrict__ T, const double4 *__restrict__ rlist,
		const int2 *__restrict__ bonds, long long *__restrict__ pstep, const float4 *__restrict__ blist) {
	const double hdt = 0.5 * S.dt;

	const FT f = F[i];
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

	const FT t = T[i];
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
	const float4 bp = write_mirror(i, r, q, S, bonds[i], q32, bpos, back64);
	pstep[i] += 1;

	double4 rl = rlist[i];
	double dx = r.x - rl.x, dy = r.y - rl.y, dz = r.z - rl.z;
	if(S.far_back && blist && back_moved(bp, blist[i], S)) return true;
	return dx * dx + dy * dy + dz * dz > S.skin2;
}

// second half kick of particle i and the Langevin / Brownian thermostat
// kick (src/CUDA/Backends/CUDA_mixed.cuh:71-93)
template<class FT>
__device__ __forceinline__ void second_step_particle(int i, const oxh_sys &S, double4 *__restrict__ vel, double4 *__restrict__ Lm,
		const FT *__restrict__ F, const FT *__restrict__ T, const int *__restrict__ orig, const long long *__restrict__ pstep) {
	const double hdt = 0.5 * S.dt;
	const FT f = F[i], t = T[i];
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
}

// ------------------------------------------------------------------ Bussi
// Gamma(a, 1) variate, a >= 1, after G. Marsaglia and W. W. Tsang,
// "A simple method for generating gamma variables", ACM TOMS 26 (2000).
__device__ double gamma_mt(const oxh_sys &S, long long step, uint32_t stream, double a) {
	double d = a - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
#pragma unroll 1
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

// sum of n squared standard normals (chi-squared with n degrees of
// freedom): twice a Gamma(n/2) variate, plus one squared normal when n is
// odd (one sampler call site keeps the register footprint small)
__device__ double sum_noises(const oxh_sys &S, long long step, uint32_t stream, int n) {
	if(n <= 0) return 0.;
	double extra = 0.;
	if(n % 2 == 1) {
		double a, b;
		uint4 w = rng(S, 0xFFFE0000u, step, stream);
		box_muller(w.x, w.y, a, b);
		extra = a * a;
	}
	if(n == 1) return extra;
	return 2.0 * gamma_mt(S, step, stream, n / 2) + extra;
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

// Sums over particles [begin, end) of K_t, K_r (1/2 v^2, 1/2 L^2) and the
// velocity, in a fixed 256-wide tree: thread t < 256 accumulates particles
// begin + t, begin + t + 256, ...; the result is left in sh[k][0].  Every
// thread of the block calls it (threads >= 256 only join the barriers).
__device__ __forceinline__ void bussi_chunk_sums(int begin, int end, const double4 *__restrict__ vel, const double4 *__restrict__ Lm,
		double (*sh)[256]) {
	const int tid = threadIdx.x;
	if(tid < 256) {
		double acc[5] = {0, 0, 0, 0, 0};
		for(int i = begin + tid; i < end; i += 256) {
			double4 v = vel[i], L = Lm[i];
			acc[0] += 0.5 * (v.x * v.x + v.y * v.y + v.z * v.z);
			acc[1] += 0.5 * (L.x * L.x + L.y * L.y + L.z * L.z);
			acc[2] += v.x;
			acc[3] += v.y;
			acc[4] += v.z;
		}
		for(int k = 0; k < 5; k++) sh[k][tid] = acc[k];
	}
	__syncthreads();
	for(int off = 128; off > 0; off >>= 1) {
		if(tid < off)
			for(int k = 0; k < 5; k++) sh[k][tid] += sh[k][tid + off];
		__syncthreads();
	}
}

// Advances the system's thermostat kinetic energies from the sums r[5] of
// bussi_chunk_sums and returns the rescaling factors of v about the
// centre-of-mass velocity vc and of L (src/Backends/Thermostats/
// BussiThermostat.cpp:54-90)
__device__ __forceinline__ void bussi_update(const oxh_sys &S, oxh_sys_state &st, long long step, const double r[5], double &ft, double &fr,
		double vc[3]) {
	const double n = S.n;
	vc[0] = r[2] / n;
	vc[1] = r[3] / n;
	vc[2] = r[4] / n;
	const double K_now_t = r[0] - 0.5 * n * (vc[0] * vc[0] + vc[1] * vc[1] + vc[2] * vc[2]);
	const double K_now_r = r[1];
	const int dof_t = 3 * (S.n - 1), dof_r = 3 * S.n;
	// translational then rotational target, through one call site
#pragma unroll 1
	for(int c = 0; c < 2; c++) {
		const double K = bussi_K(S, step, RNG_BUSSI + 4 * c, c ? st.K_r : st.K_t, c ? dof_r : dof_t);
		if(c) st.K_r = K;
		else st.K_t = K;
	}
	ft = K_now_t > 0. ? sqrt(st.K_t / K_now_t) : 1.0;
	fr = K_now_r > 0. ? sqrt(st.K_r / K_now_r) : 1.0;
}

// v <- vc + (v - vc) ft, L <- L fr
__device__ __forceinline__ void bussi_rescale_particle(int i, double4 *__restrict__ vel, double4 *__restrict__ Lm, double ft, double fr,
		double vcx, double vcy, double vcz) {
	double4 v = vel[i];
	v.x = (v.x - vcx) * ft + vcx;
	v.y = (v.y - vcy) * ft + vcy;
	v.z = (v.z - vcz) * ft + vcz;
	vel[i] = v;
	double4 L = Lm[i];
	L.x *= fr;
	L.y *= fr;
	L.z *= fr;
	Lm[i] = L;
}

}  // namespace oxh
