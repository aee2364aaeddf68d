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
// (src/Observables/PotentialEnergy.cpp:21-33).  The formulas are the CPU
// classes' force() methods: where the reference's CUDA kernel differs from
// its CPU class (the ellipsoid's extra sphere term, the cone's angle form,
// the COM force on stored rather than absolute positions), the CPU class
// is followed, as it is the one the reference's own tests run.
//
// "com" forces share their centres through groups: one workgroup per group
// sums the absolute positions of its com and reference lists in FP64
// before the per-particle kernel reads them, instead of every particle of
// the group summing both lists itself.
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

__device__ __forceinline__ double dot3(double3 a, const double *v) {
	return a.x * v[0] + a.y * v[1] + a.z * v[2];
}

__device__ __forceinline__ double dotd(double3 a, double3 b) {
	return a.x * b.x + a.y * b.y + a.z * b.z;
}

__device__ __forceinline__ double3 scale3(double3 a, double f) {
	return make_double3(a.x * f, a.y * f, a.z * f);
}

__device__ __forceinline__ double3 cross3(double3 a, double3 b) {
	return make_double3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// LR_vector::normalize (src/Utilities/LR_vector.h:86-89,131-136): unchanged
// when zero, else times the reciprocal of the module
__device__ __forceinline__ double3 normalized(double3 a) {
	const double n2 = dotd(a, a);
	if(n2 == 0.) return a;
	return scale3(a, 1. / sqrt(n2));
}

// the star metric of RepulsiveKeplerPoinsot (star_metric_and_grad,
// src/Forces/RepulsiveKeplerPoinsot.cpp:105-226): over the 12 icosahedral
// spike axes whose span [base, apex] holds p's projection, the smallest
// ratio of the soft-max pentagon support of p's perpendicular part to the
// spike's radius there, and its gradient
__device__ bool star_metric(double3 p, double base, double apex, double base_radius, double kappa, double &s_best, double3 &grad_best) {
	const double phi = (1.0 + sqrt(5.0)) * 0.5;
	const double ax[12][3] = {{0, 1, phi}, {0, -1, phi}, {0, 1, -phi}, {0, -1, -phi}, {1, phi, 0}, {-1, phi, 0}, {1, -phi, 0},
			{-1, -phi, 0}, {phi, 0, 1}, {-phi, 0, 1}, {phi, 0, -1}, {-phi, 0, -1}};
	double cth[5], sth[5];
	for(int j = 0; j < 5; j++) {
		const double th = (2.0 * M_PI) * (double) j / 5.0;
		cth[j] = cos(th);
		sth[j] = sin(th);
	}
	bool found = false;
	s_best = 1e30;
	grad_best = make_double3(0., 0., 0.);
	const double denom = apex - base;
	for(int i = 0; i < 12; i++) {
		// the reference's _normalize divides by the norm
		double3 n = make_double3(ax[i][0], ax[i][1], ax[i][2]);
		const double nn = sqrt(dotd(n, n));
		n = make_double3(n.x / nn, n.y / nn, n.z / nn);
		const double t = dotd(n, p);
		if(t < base || t > apex || denom <= 0.) continue;
		const double R = base_radius * (apex - t) / denom;
		if(R <= 1e-12) continue;
		const double3 a = fabs(n.z) < 0.9 ? make_double3(0., 0., 1.) : make_double3(0., 1., 0.);
		double3 u = cross3(n, a);
		const double un = sqrt(dotd(u, u));
		if(un <= 0.) continue;
		u = make_double3(u.x / un, u.y / un, u.z / un);
		if(sqrt(dotd(u, u)) <= 1e-12) continue;
		const double3 v = cross3(n, u);
		const double3 q = make_double3(p.x - t * n.x, p.y - t * n.y, p.z - t * n.z);
		const double x = dotd(q, u), y = dotd(q, v);
		double proj[5], m = -1e30;
		for(int j = 0; j < 5; j++) {
			proj[j] = x * cth[j] + y * sth[j];
			if(proj[j] > m) m = proj[j];
		}
		double w[5], sumw = 0.;
		for(int j = 0; j < 5; j++) {
			w[j] = exp(kappa * (proj[j] - m));
			sumw += w[j];
		}
		if(sumw <= 0.) continue;
		const double support = m + (1.0 / kappa) * log(sumw);
		double dx = 0., dy = 0.;
		for(int j = 0; j < 5; j++) {
			const double sj = w[j] / sumw;
			dx += sj * cth[j];
			dy += sj * sth[j];
		}
		const double3 gsup = make_double3(u.x * dx + v.x * dy, u.y * dx + v.y * dy, u.z * dx + v.z * dy);
		const double s = support / R;
		if(s < s_best) {
			const double Rp = -base_radius / denom;
			const double dinv = -(Rp) / (R * R);
			const double invR = 1.0 / R;
			s_best = s;
			grad_best = make_double3(invR * gsup.x + support * (dinv * n.x), invR * gsup.y + support * (dinv * n.y),
					invR * gsup.z + support * (dinv * n.z));
			found = true;
		}
	}
	return found;
}

__device__ __forceinline__ double3 ext_force(const oxh_ext_rec &r, long long step, double3 p, const oxh_sys &S, const double4 *__restrict__ pos,
		const int4 *__restrict__ img, const int *__restrict__ slot_of, const double *__restrict__ com, int &err) {
	const double st = (double) step;
	const double3 zero = make_double3(0., 0., 0.);
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
	case OXH_EXT_LOWDIM_TRAP: {
		// LowdimMovingTrap (src/Forces/LowdimMovingTrap.cpp:67-81): only the
		// visible components (flags bits 0-2) pull
		const double s = r.b * st;
		return make_double3((r.flags & 1) ? -r.a * (p.x - (r.v2[0] + s * r.v1[0])) : 0.,
				(r.flags & 2) ? -r.a * (p.y - (r.v2[1] + s * r.v1[1])) : 0., (r.flags & 4) ? -r.a * (p.z - (r.v2[2] + s * r.v1[2])) : 0.);
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
		if(d >= 0.) return zero;
		const double f = -(d * r.a);
		return make_double3(f * r.v1[0], f * r.v1[1], f * r.v1[2]);
	}
	case OXH_EXT_ATTRACTION_PLANE: {
		// AttractionPlane (src/Forces/AttractionPlane.cpp:43-51): a
		// constant pull above the plane, harmonic below it (a = stiff,
		// b = position)
		const double d = r.v1[0] * p.x + r.v1[1] * p.y + r.v1[2] * p.z + r.b;
		const double f = d >= 0. ? -r.a * 1.0 : -(d * r.a);
		return make_double3(f * r.v1[0], f * r.v1[1], f * r.v1[2]);
	}
	case OXH_EXT_PLANE_MOVING: {
		// RepulsionPlaneMoving (src/Forces/RepulsionPlaneMoving.cpp:48-55):
		// one plane through each reference particle ref..ref2
		double3 f = zero;
		const double kx = -r.a * r.v1[0], ky = -r.a * r.v1[1], kz = -r.a * r.v1[2];
		for(int j = r.ref; j <= r.ref2; j++) {
			const double3 q = abs_pos(pos, img, slot_of[S.first + j], S);
			const double d = fmin((p.x - q.x) * r.v1[0] + (p.y - q.y) * r.v1[1] + (p.z - q.z) * r.v1[2], 0.);
			f.x += kx * d;
			f.y += ky * d;
			f.z += kz * d;
		}
		return f;
	}
	case OXH_EXT_SPHERE: {
		// RepulsiveSphere (src/Forces/RepulsiveSphere.cpp:50-58)
		const double3 dist = min_image(make_double3(r.v1[0], r.v1[1], r.v1[2]), p, S);
		const double m = sqrt(dist.x * dist.x + dist.y * dist.y + dist.z * dist.z);
		const double radius = r.c + r.b * st;
		if(m <= radius || m >= r.d) return zero;
		const double f = -r.a * (1. - radius / m);
		return make_double3(dist.x * f, dist.y * f, dist.z * f);
	}
	case OXH_EXT_SPHERE_SMOOTH: {
		// RepulsiveSphereSmooth (src/Forces/RepulsiveSphereSmooth.cpp:
		// 50-63): a = stiff, c = r0, d = r_ext, e = smooth, f = alpha
		const double3 dist = min_image(make_double3(r.v1[0], r.v1[1], r.v1[2]), p, S);
		const double m = sqrt(dotd(dist, dist));
		if(m < r.c || m > r.d) return zero;
		if(m >= r.f && m <= r.d) return scale3(dist, -(r.a * 0.5 * exp((m - r.f) / r.e)) / m);
		return scale3(dist, -(r.a * m - r.a * 0.5 * exp(-(m - r.f) / r.e)) / m);
	}
	case OXH_EXT_SPHERE_MOVING: {
		// RepulsiveSphereMoving (src/Forces/RepulsiveSphereMoving.cpp:
		// 89-148): a WCA wall (2x-x form with x = 2, sigma = 1, epsilon =
		// stiff) on the gap to a sphere whose centre moves from v1 to v2
		// over `steps` steps and whose radius c grows at rate b; d = r_ext
		double3 c = make_double3(r.v1[0], r.v1[1], r.v1[2]);
		if(r.steps > 0) {
			double t = st / (double) r.steps;
			if(t < 0.) t = 0.;
			if(t > 1.) t = 1.;
			c = make_double3(c.x + (r.v2[0] - c.x) * t, c.y + (r.v2[1] - c.y) * t, c.z + (r.v2[2] - c.z) * t);
		}
		const double3 dist = min_image(c, p, S);
		const double m = sqrt(dotd(dist, dist));
		const double gap = m - (r.c + r.b * st);
		if(gap >= r.d || m <= 0.0 || gap >= sqrt(2.0)) return zero;
		const double rs = gap > 1e-9 ? gap : 1e-9;
		const double y = 1.0 / rs, A = y * y;
		const double dUdr = 4.0 * r.a * (2.0 * A - 1.0) * (-(2.0 / rs) * A);
		return scale3(dist, -dUdr / m);
	}
	case OXH_EXT_KEPLER_POINSOT: {
		// RepulsiveKeplerPoinsot (src/Forces/RepulsiveKeplerPoinsot.cpp:
		// 228-268): a WCA wall (x = 4) in the star metric s, acting inside
		// s < 1; the shape (c = apex, d = base, e = base_radius) grows as
		// 1 + rate * step; f = kappa
		const double growth = 1.0 + r.b * st;
		if(growth <= 0.) return zero;
		const double base = r.d * growth, apex = r.c * growth, base_radius = r.e * growth;
		if(apex <= base || base_radius <= 0.) return zero;
		const double3 q = min_image(make_double3(r.v1[0], r.v1[1], r.v1[2]), p, S);
		double s;
		double3 g;
		if(!star_metric(q, base, apex, base_radius, r.f, s, g) || s >= 1.0) return zero;
		const double ss = s > 1e-9 ? s : 1e-9;
		const double y = 1.0 / ss, A = (y * y) * (y * y);
		const double dUds = 4.0 * r.a * (2.0 * A - 1.0) * (-(4.0 / ss) * A);
		return scale3(g, -dUds);
	}
	case OXH_EXT_LJ_WALL: {
		// LJWall (src/Forces/LJWall.cpp:59-65): a = stiff, b = position,
		// c = sigma, d = cutoff (in units of sigma), n
		const double d = dot3(p, r.v1) + r.b;
		const double rel = d / r.c;
		if(rel > r.d) return zero;
		const double lj = pow(rel, (double) -r.n);
		const double f = 4 * r.n * r.a * (2 * (lj * lj) - lj) / d;
		return make_double3(r.v1[0] * f, r.v1[1] * f, r.v1[2] * f);
	}
	case OXH_EXT_TWIST: {
		// ConstantRateTorque (src/Forces/ConstantRateTorque.cpp:107-140): a
		// trap at pos0 (v2) turned by base + rate * step radians (c + b *
		// step) about the axis v1 through the centre v3, its components
		// masked by v4
		const double t = r.c + r.b * st;
		const double sn = sin(t), cs = cos(t), oc = 1. - cs;
		const double *u = r.v1;
		const double xyo = u[0] * u[1] * oc, xzo = u[0] * u[2] * oc, yzo = u[1] * u[2] * oc;
		const double xs = u[0] * sn, ys = u[1] * sn, zs = u[2] * sn;
		const double rx = r.v2[0] - r.v3[0], ry = r.v2[1] - r.v3[1], rz = r.v2[2] - r.v3[2];
		const double tx = (u[0] * u[0] * oc + cs) * rx + (xyo - zs) * ry + (xzo + ys) * rz + r.v3[0];
		const double ty = (xyo + zs) * rx + (u[1] * u[1] * oc + cs) * ry + (yzo - xs) * rz + r.v3[1];
		const double tz = (xzo - ys) * rx + (yzo + xs) * ry + (u[2] * u[2] * oc + cs) * rz + r.v3[2];
		return make_double3(-r.a * (p.x - tx) * r.v4[0], -r.a * (p.y - ty) * r.v4[1], -r.a * (p.z - tz) * r.v4[2]);
	}
	case OXH_EXT_CENTRAL: {
		// GenericCentralForce with force_type = gravity (src/Forces/
		// GenericCentralForce.cpp:176-197): F0 (a) towards the centre v1
		// between the cut-offs c (inner) and d (outer; 0 = none)
		const double3 dir = make_double3(r.v1[0] - p.x, r.v1[1] - p.y, r.v1[2] - p.z);
		const double d2 = dotd(dir, dir);
		if(d2 < r.c * r.c) return zero;
		if(r.d > 0. && d2 > r.d * r.d) return zero;
		return scale3(normalized(dir), r.a);
	}
	case OXH_EXT_LJ_CONE: {
		// LJCone (src/Forces/LJCone.cpp:64-87): a = stiff, c = sigma,
		// d = cutoff, e, f, g = sine, cosine and tangent of the half-angle,
		// v1 = axis, v2 = apex
		const double3 va = make_double3(p.x - r.v2[0], p.y - r.v2[1], p.z - r.v2[2]);
		const double dal = dot3(va, r.v1);
		const double3 vax = make_double3(r.v1[0] * dal - va.x, r.v1[1] * dal - va.y, r.v1[2] * dal - va.z);
		const double dax = sqrt(dotd(vax, vax));
		const double dcone = dal * r.e - dax * r.f;
		const double rel = dcone / r.c;
		if(rel > r.d) return zero;
		const double k = dal + dax * r.g;
		const double3 nrm = normalized(make_double3(k * r.v1[0] - va.x, k * r.v1[1] - va.y, k * r.v1[2] - va.z));
		const double lj = pow(rel, (double) -r.n);
		return scale3(nrm, 4 * r.n * r.a * (2 * (lj * lj) - lj) / dcone);
	}
	case OXH_EXT_ELLIPSOID: {
		// RepulsiveEllipsoid (src/Forces/RepulsiveEllipsoid.cpp:51-61): a
		// constant push (a) towards the centre v1 outside the inner
		// ellipsoid (semi-axes r_2 = v3) or inside the outer one (r_1 = v2)
		const double3 dist = min_image(make_double3(r.v1[0], r.v1[1], r.v1[2]), p, S);
		const double in = dist.x * dist.x / (r.v3[0] * r.v3[0]) + dist.y * dist.y / (r.v3[1] * r.v3[1]) + dist.z * dist.z / (r.v3[2] * r.v3[2]);
		const double ex = dist.x * dist.x / (r.v2[0] * r.v2[0]) + dist.y * dist.y / (r.v2[1] * r.v2[1]) + dist.z * dist.z / (r.v2[2] * r.v2[2]);
		if(in < 1. && ex > 1.) return zero;
		const double3 u = normalized(dist);
		return make_double3(-r.a * u.x, -r.a * u.y, -r.a * u.z);
	}
	case OXH_EXT_COM: {
		// COMForce (src/Forces/COMForce.cpp:57-64): a spring between the
		// centre of the com list and that of the reference list, shared
		// over the n com particles (a = stiff, b = rate, c = r0; the centres
		// of group ref)
		const double *g = com + 6 * r.ref;
		const double3 dist = make_double3(g[3] - g[0], g[4] - g[1], g[5] - g[2]);
		const double dc = sqrt(dotd(dist, dist));
		const double f = (dc - (r.c + r.b * st)) * r.a / (double) r.n;
		return scale3(dist, f / dc);
	}
	case OXH_EXT_YUKAWA_SPHERE: {
		// YukawaSphere (src/Forces/YukawaSphere.cpp:56-77): a = radius,
		// b = WCA epsilon, c = WCA sigma, d = Debye length, e = Debye
		// amplitude, f = cutoff, g = WCA cutoff, n = WCA n; a particle
		// beyond the radius is an error there
		const double3 dist = min_image(make_double3(r.v1[0], r.v1[1], r.v1[2]), p, S);
		const double m = sqrt(dotd(dist, dist));
		const double ds = r.a - m;
		if(ds < 0.) {
			err |= ERR_EXT;
			return zero;
		}
		if(ds >= r.f) return zero;
		const double3 dir = scale3(dist, -(1. / m));
		double3 f = scale3(dir, (r.e * exp(-ds / r.d)) * (1.0 / (ds * r.d) + 1.0 / (ds * ds)));
		if(ds < r.g) {
			const double w = pow(r.c / ds, 6.);
			const double h = 4 * r.b * r.n * (2 * (w * w) - w) / ds;
			f.x += dir.x * h;
			f.y += dir.y * h;
			f.z += dir.z * h;
		}
		return f;
	}
	default: return zero;
	}
}

// one workgroup per "com" group: the centres of its com list (com[0..2])
// and of its reference list (com[3..5]), from absolute positions
__global__ void __launch_bounds__(256) k_ext_groups(int n_grp, const int4 *__restrict__ grp, const int *__restrict__ grp_sys,
		const int *__restrict__ idx, const double4 *__restrict__ pos, const int4 *__restrict__ img, const int *__restrict__ slot_of,
		const oxh_sys *__restrict__ systems, double *__restrict__ com) {
	__shared__ double red[6][256];
	const int g = blockIdx.x;
	const oxh_sys &S = systems[grp_sys[g]];
	const int4 G = grp[g];
	double acc[6] = {0., 0., 0., 0., 0., 0.};
	for(int k = threadIdx.x; k < G.y; k += blockDim.x) {
		const double3 q = abs_pos(pos, img, slot_of[S.first + idx[G.x + k]], S);
		acc[0] += q.x;
		acc[1] += q.y;
		acc[2] += q.z;
	}
	for(int k = threadIdx.x; k < G.w; k += blockDim.x) {
		const double3 q = abs_pos(pos, img, slot_of[S.first + idx[G.z + k]], S);
		acc[3] += q.x;
		acc[4] += q.y;
		acc[5] += q.z;
	}
	for(int c = 0; c < 6; c++) red[c][threadIdx.x] = acc[c];
	__syncthreads();
	for(int w = blockDim.x / 2; w > 0; w >>= 1) {
		if((int) threadIdx.x < w)
			for(int c = 0; c < 6; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
		__syncthreads();
	}
	if(threadIdx.x < 6) {
		const int c = threadIdx.x;
		com[6 * g + c] = red[c][0] * (1. / (double) (c < 3 ? G.y : G.w));
	}
}

template<class FT>
__global__ void k_ext_forces(int N, const double4 *__restrict__ pos, const int4 *__restrict__ img, const int *__restrict__ orig,
		const int2 *__restrict__ pinfo, const oxh_sys *__restrict__ systems, const long long *__restrict__ pstep,
		const int *__restrict__ slot_of, const int *__restrict__ ext_off, const oxh_ext_rec *__restrict__ rec, const double *__restrict__ com,
		oxh_sys_state *__restrict__ state, FT *__restrict__ F) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= N) return;
	const int s = pinfo[i].y;
	const oxh_sys &S = systems[s];
	if(!S.has_ext || i >= S.first + S.n) return;
	const int o = S.first + orig[i];
	const int r0 = ext_off[o], r1 = ext_off[o + 1];
	if(r0 == r1) return;
	const long long step = S.ext_step0 + (pstep[i] > 0 ? pstep[i] - 1 : 0);
	const double3 p = abs_pos(pos, img, i, S);
	double fx = 0., fy = 0., fz = 0.;
	int err = 0;
	for(int k = r0; k < r1; k++) {
		const double3 f = ext_force(rec[k], step, p, S, pos, img, slot_of, com, err);
		fx += f.x;
		fy += f.y;
		fz += f.z;
	}
	if(err) atomicOr(&state[s].error, err);
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
	if(b.n_grp > 0)
		k_ext_groups<<<b.n_grp, 256, 0, st>>>(b.n_grp, b.ext_grp, b.ext_grp_sys, b.ext_idx, b.pos, b.img, b.slot_of, b.d_sys, b.ext_com);
	const int bs = 256, g = (b.N + bs - 1) / bs;
	if(b.fp64)
		k_ext_forces<double4><<<g, bs, 0, st>>>(b.N, b.pos, b.img, b.orig, b.pinfo, b.d_sys, b.pstep, b.slot_of, b.ext_off, b.ext_rec,
				b.ext_com, b.d_state, b.F64);
	else
		k_ext_forces<float4><<<g, bs, 0, st>>>(b.N, b.pos, b.img, b.orig, b.pinfo, b.d_sys, b.pstep, b.slot_of, b.ext_off, b.ext_rec,
				b.ext_com, b.d_state, b.F);
}

void launch_slot_of(hipStream_t st, const Batch &b) {
	if(b.N == 0) return;
	k_slot_of<<<(b.N + 255) / 256, 256, 0, st>>>(b.N, b.orig, b.pinfo, b.d_sys, b.slot_of);
}

}  // namespace oxh
