// External forces on gfx950 (the reference's external_forces_file:
// src/Forces/*.cpp on the CPU, the CUDA_trap family of
// src/CUDA/Backends/CUDA_MD.cuh:97-554 on its GPU).
//
// Records (oxh_ext_rec) are stored per particle in CSR form, keyed by the
// particle's index inside its system as uploaded, so that a Hilbert sort
// changes nothing here: ext_off[first + o] .. ext_off[first + o + 1] are the
// records of particle o of the system that starts at `first`.  One thread per
// particle slot evaluates its records in order in FP64 and adds the sum to
// the force the pair kernels wrote (the reference sets F to the external
// forces and lets the interaction accumulate onto it; sums here run in the
// other order, which only moves the last bits).  None of these types has a
// torque or enters the potential energy the energy file reports
// (src/Observables/PotentialEnergy.cpp:21-33).
//
// Positions are the absolute ones the reference passes to every force
// (BaseBox::get_abs_pos: the stored position plus its image count times the
// box, src/Boxes/OrthogonalBox.cpp:63-65), and the step is the reference's
// step counter of the force evaluation (ext_step0 + the particle's step
// index).
#include "../../include/oxh.h"
#include "oxh_device.h"
#include "oxh_internal.h"

namespace oxh {

namespace {

__device__ __forceinline__ double3 abs_pos(const double4 *__restrict__ pos, const int4 *__restrict__ img, int i, const oxh_sys &S) {
	const double4 p = pos[i];
	const int4 m = img[i];
	return make_double3(p.x + S.box[0] * (double) m.x, p.y + S.box[1] * (double) m.y, p.z + S.box[2] * (double) m.z);
}

// nearest image of v - u in an orthogonal box (CubicBox / OrthogonalBox::
// min_image, src/Boxes/OrthogonalBox.cpp)
__device__ __forceinline__ double3 min_image(double3 u, double3 v, const oxh_sys &S) {
	double dx = v.x - u.x, dy = v.y - u.y, dz = v.z - u.z;
	dx -= S.box[0] * rint(dx / S.box[0]);
	dy -= S.box[1] * rint(dy / S.box[1]);
	dz -= S.box[2] * rint(dz / S.box[2]);
	return make_double3(dx, dy, dz);
}

__device__ __forceinline__ double3 ext_force(const oxh_ext_rec &r, long long step, double3 p, const oxh_sys &S, const double4 *__restrict__ pos,
		const int4 *__restrict__ img, const int *__restrict__ slot_of) {
	const double st = (double) step;
	switch(r.type) {
	case OXH_EXT_STRING: {
		// ConstantRateForce (src/Forces/ConstantRateForce.cpp:44-53)
		double dx = r.v1[0], dy = r.v1[1], dz = r.v1[2];
		if(r.flags & 2) {
			dx -= p.x;
			dy -= p.y;
			dz -= p.z;
			const double m = sqrt(dx * dx + dy * dy + dz * dz);
			dx /= m;
			dy /= m;
			dz /= m;
		}
		const double f = r.a + r.b * st;
		return make_double3(f * dx, f * dy, f * dz);
	}
	case OXH_EXT_TRAP: {
		// MovingTrap (src/Forces/MovingTrap.cpp:55-68)
		const double s = r.b * st;
		return make_double3(-r.a * (p.x - (r.v2[0] + s * r.v1[0])), -r.a * (p.y - (r.v2[1] + s * r.v1[1])),
				-r.a * (p.z - (r.v2[2] + s * r.v1[2])));
	}
	case OXH_EXT_MUTUAL_TRAP:
	case OXH_EXT_CONSTANT_TRAP: {
		// MutualTrap (src/Forces/MutualTrap.cpp:62-76), ConstantTrap
		// (src/Forces/ConstantTrap.cpp:49-58): dr = partner - self
		const double3 q = abs_pos(pos, img, slot_of[S.first + r.ref], S);
		double3 dr = (r.flags & 1) ? min_image(p, q, S) : make_double3(q.x - p.x, q.y - p.y, q.z - p.z);
		const double m = sqrt(dr.x * dr.x + dr.y * dr.y + dr.z * dr.z);
		double f;
		if(r.type == OXH_EXT_MUTUAL_TRAP) f = (m - (r.c + r.d * st)) * (r.a + r.b * st) / m;
		else f = r.a * copysign(1., m - r.c) / m;
		return make_double3(dr.x * f, dr.y * f, dr.z * f);
	}
	case OXH_EXT_REPULSION_PLANE: {
		// RepulsionPlane (src/Forces/RepulsionPlane.cpp:56-74)
		double position = r.b + r.d * st;
		if(r.c > r.b && position > r.c) position = r.c;
		if(r.c < r.b && position < r.c) position = r.c;
		const double d = r.v1[0] * p.x + r.v1[1] * p.y + r.v1[2] * p.z + position;
		if(d >= 0.) return make_double3(0., 0., 0.);
		const double f = -(d * r.a);
		return make_double3(f * r.v1[0], f * r.v1[1], f * r.v1[2]);
	}
	case OXH_EXT_SPHERE: {
		// RepulsiveSphere (src/Forces/RepulsiveSphere.cpp:50-58)
		const double3 dist = min_image(make_double3(r.v1[0], r.v1[1], r.v1[2]), p, S);
		const double m = sqrt(dist.x * dist.x + dist.y * dist.y + dist.z * dist.z);
		const double radius = r.c + r.b * st;
		if(m <= radius || m >= r.d) return make_double3(0., 0., 0.);
		const double f = -r.a * (1. - radius / m);
		return make_double3(dist.x * f, dist.y * f, dist.z * f);
	}
	default: return make_double3(0., 0., 0.);
	}
}

template<class FT>
__global__ void k_ext_forces(int N, const double4 *__restrict__ pos, const int4 *__restrict__ img, const int *__restrict__ orig,
		const int2 *__restrict__ pinfo, const oxh_sys *__restrict__ systems, const long long *__restrict__ pstep,
		const int *__restrict__ slot_of, const int *__restrict__ ext_off, const oxh_ext_rec *__restrict__ rec, FT *__restrict__ F) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= N) return;
	const oxh_sys &S = systems[pinfo[i].y];
	if(!S.has_ext || i >= S.first + S.n) return;
	const int o = S.first + orig[i];
	const int r0 = ext_off[o], r1 = ext_off[o + 1];
	if(r0 == r1) return;
	const long long step = S.ext_step0 + (pstep[i] > 0 ? pstep[i] - 1 : 0);
	const double3 p = abs_pos(pos, img, i, S);
	double fx = 0., fy = 0., fz = 0.;
	for(int k = r0; k < r1; k++) {
		const double3 f = ext_force(rec[k], step, p, S, pos, img, slot_of);
		fx += f.x;
		fy += f.y;
		fz += f.z;
	}
	FT v = F[i];
	v.x += fx;
	v.y += fy;
	v.z += fz;
	F[i] = v;
}

__global__ void k_slot_of(int N, const int *__restrict__ orig, const int2 *__restrict__ pinfo, const oxh_sys *__restrict__ systems,
		int *__restrict__ slot_of) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= N) return;
	const oxh_sys &S = systems[pinfo[i].y];
	if(i >= S.first + S.n) return;
	slot_of[S.first + orig[i]] = i;
}

}  // namespace

void launch_ext_forces(hipStream_t st, const Batch &b) {
	if(b.n_ext == 0 || b.N == 0) return;
	const int bs = 256, g = (b.N + bs - 1) / bs;
	if(b.fp64)
		k_ext_forces<double4><<<g, bs, 0, st>>>(b.N, b.pos, b.img, b.orig, b.pinfo, b.d_sys, b.pstep, b.slot_of, b.ext_off, b.ext_rec, b.F64);
	else
		k_ext_forces<float4><<<g, bs, 0, st>>>(b.N, b.pos, b.img, b.orig, b.pinfo, b.d_sys, b.pstep, b.slot_of, b.ext_off, b.ext_rec, b.F);
}

void launch_slot_of(hipStream_t st, const Batch &b) {
	if(b.N == 0) return;
	k_slot_of<<<(b.N + 255) / 256, 256, 0, st>>>(b.N, b.orig, b.pinfo, b.d_sys, b.slot_of);
}

}  // namespace oxh
