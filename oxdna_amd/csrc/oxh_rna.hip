// oxRNA / oxRNA2 pair forces on gfx950, every term in FP64.
//
// Terms: src/Interactions/RNAInteraction.cpp:433-1153 (backbone FENE,
// bonded excluded volume, stacking, non-bonded excluded volume, hydrogen
// bonding, cross stacking, coaxial stacking), :1195-1219 (repulsive LJ),
// :1221-1407 (f1, f2, f4, f5 and their derivatives) and the oxRNA2
// Debye-Hueckel term (src/Interactions/RNAInteraction2.cpp:135-185).  The
// hydrogen-bonding f4 factors are analytic, as in the reference's GPU
// kernel (src/CUDA/Interactions/CUDA_RNA.cuh:743-761).  Interaction sites:
// src/Particles/RNANucleotide.h:25-47.
//
// One kernel evaluates a nucleotide's bonded neighbours and its whole
// Verlet row (close pairs: every term; far pairs, centres beyond the
// close bound: Debye-Hueckel only) with K lanes per nucleotide, as the
// oxDNA2 kernels do, and writes (not accumulates) F and T.  The mixed form
// takes the angles of the angular factors (acos, sin, f4) in FP32, every
// other quantity in FP64, and rounds the results into its FP32 force
// buffers; the all-FP64 form takes everything in FP64.
#include <cfloat>
#include <type_traits>

#include "../../include/oxh.h"
#include "oxh_internal.h"
#include "oxh_device.h"
#include "oxh_rna.h"

namespace oxh {
namespace rna {

// ------------------------------------------------ modulation functions
// The tables of the reference hold the float model constants as doubles;
// every expression below promotes them the same way.  The f4 / f5 type is
// a template argument, so that each table entry folds into a constant.

// the angle of a cosine, clamped as the reference's LRACOS (src/defs.h),
// in the angle type A: double in the all-FP64 form, float in the mixed one
template<typename A>
__device__ __forceinline__ A lracos(double x) { return x > 1. ? A(0) : (x < -1. ? (A) PIf : acos((A) x)); }

template<int TY>
__device__ __forceinline__ double f1(const oxh_sys &S, double r, int n3, int n5) {
	double val = 0.;
	if(r < (double) F1_RCHIGH[TY]) {
		if(r > (double) F1_RHIGH[TY]) val = S.f1_eps_d[TY][n3][n5] * (double) F1_BHIGH[TY] * SQR(r - (double) F1_RCHIGH[TY]);
		else if(r > (double) F1_RLOW[TY]) {
			const double tmp = 1 - exp(-(r - (double) F1_R0[TY]) * (double) F1_A[TY]);
			val = S.f1_eps_d[TY][n3][n5] * SQR(tmp) - S.f1_shift_d[TY][n3][n5];
		}
		else if(r > (double) F1_RCLOW[TY]) val = S.f1_eps_d[TY][n3][n5] * (double) F1_BLOW[TY] * SQR(r - (double) F1_RCLOW[TY]);
	}
	return val;
}

template<int TY>
__device__ __forceinline__ double f1D(const oxh_sys &S, double r, int n3, int n5) {
	double val = 0.;
	if(r < (double) F1_RCHIGH[TY]) {
		if(r > (double) F1_RHIGH[TY]) val = 2 * (double) F1_BHIGH[TY] * (r - (double) F1_RCHIGH[TY]);
		else if(r > (double) F1_RLOW[TY]) {
			const double tmp = exp(-(r - (double) F1_R0[TY]) * (double) F1_A[TY]);
			val = 2 * (1 - tmp) * tmp * (double) F1_A[TY];
		}
		else if(r > (double) F1_RCLOW[TY]) val = 2 * (double) F1_BLOW[TY] * (r - (double) F1_RCLOW[TY]);
	}
	return S.f1_eps_d[TY][n3][n5] * val;
}

template<int TY>
__device__ __forceinline__ double f2(double r) {
	double val = 0.;
	if(r < (double) F2_RCHIGH[TY]) {
		if(r > (double) F2_RHIGH[TY]) val = (double) F2_K[TY] * (double) F2_BHIGH[TY] * SQR(r - (double) F2_RCHIGH[TY]);
		else if(r > (double) F2_RLOW[TY])
			val = ((double) F2_K[TY] / 2.) * (SQR(r - (double) F2_R0[TY]) - SQR((double) F2_RC[TY] - (double) F2_R0[TY]));
		else if(r > (double) F2_RCLOW[TY]) val = (double) F2_K[TY] * (double) F2_BLOW[TY] * SQR(r - (double) F2_RCLOW[TY]);
	}
	return val;
}

template<int TY>
__device__ __forceinline__ double f2D(double r) {
	double val = 0.;
	if(r < (double) F2_RCHIGH[TY]) {
		if(r > (double) F2_RHIGH[TY]) val = 2. * (double) F2_K[TY] * (double) F2_BHIGH[TY] * (r - (double) F2_RCHIGH[TY]);
		else if(r > (double) F2_RLOW[TY]) val = (double) F2_K[TY] * (r - (double) F2_R0[TY]);
		else if(r > (double) F2_RCLOW[TY]) val = 2. * (double) F2_K[TY] * (double) F2_BLOW[TY] * (r - (double) F2_RCLOW[TY]);
	}
	return val;
}

template<int TY, typename A>
__device__ __forceinline__ A f4(A t) {
	A val = 0;
	t -= (A) F4_T0[TY];
	if(t < 0) t *= -1;
	if(t < (A) F4_TC[TY]) {
		if(t > (A) F4_TS[TY]) val = (A) F4_B[TY] * SQR((A) F4_TC[TY] - t);
		else val = A(1) - (A) F4_A[TY] * SQR(t);
	}
	return val;
}

// (df4/dtheta) / sin(theta), with the reference's guard for a small sine
template<int TY, typename A>
__device__ __forceinline__ A f4Dsin(A t) {
	A val = 0, m = 1;
	A tt0 = t - (A) F4_T0[TY];
	if(tt0 < 0) {
		tt0 *= -1;
		m = -1;
	}
	if(tt0 < (A) F4_TC[TY]) {
		const A sint = sin(t);
		if(tt0 > (A) F4_TS[TY]) val = m * 2 * (A) F4_B[TY] * (tt0 - (A) F4_TC[TY]) / sint;
		else if(SQR(sint) > A(1e-8)) val = -m * 2 * (A) F4_A[TY] * tt0 / sint;
		else val = -m * 2 * (A) F4_A[TY];
	}
	return val;
}

template<int TY>
__device__ __forceinline__ double f5(double f) {
	double val = 0.;
	if(f > (double) F5_XC[TY]) {
		if(f < (double) F5_XS[TY]) val = (double) F5_B[TY] * SQR((double) F5_XC[TY] - f);
		else if(f < 0) val = 1. - (double) F5_A[TY] * SQR(f);
		else val = 1.;
	}
	return val;
}

template<int TY>
__device__ __forceinline__ double f5D(double f) {
	double val = 0.;
	if(f > (double) F5_XC[TY]) {
		if(f < (double) F5_XS[TY]) val = 2 * (double) F5_B[TY] * (f - (double) F5_XC[TY]);
		else if(f < 0) val = -2. * (double) F5_A[TY] * f;
	}
	return val;
}

// truncated LJ with a quadratic shell; force on the second site
__device__ __forceinline__ double repulsive_lj(const dv3 &r, dv3 &force, float sigma, float rstar, float b, float rc) {
	const double rnorm = dot(r, r);
	double energy = 0.;
	force = {0., 0., 0.};
	if(rnorm < SQR((double) rc)) {
		if(rnorm > SQR((double) rstar)) {
			const double rmod = sqrt(rnorm);
			const double rrc = rmod - (double) rc;
			energy = (double) EXCL_EPS * (double) b * SQR(rrc);
			force = r * (-(2 * (double) EXCL_EPS * (double) b * rrc / rmod));
		}
		else {
			const double tmp = SQR((double) sigma) / rnorm;
			const double lj = tmp * tmp * tmp;
			energy = 4 * (double) EXCL_EPS * (SQR(lj) - lj);
			force = r * (-(24 * (double) EXCL_EPS * (lj - 2 * SQR(lj)) / rnorm));
		}
	}
	return energy;
}

// ------------------------------------------------------ site geometry
// A nucleotide's axes (columns of its orientation matrix) from its unit
// quaternion, and its base type.
struct Nuc {
	dv3 a1, a2, a3;
	int type, btype;
};

__device__ __forceinline__ int type_of(int btype) { return (btype < 0) ? 3 - ((3 - btype) % 4) : btype % 4; }

__device__ __forceinline__ Nuc nuc(double4 q, int btype) {
	const double sqx = q.x * q.x, sqy = q.y * q.y, sqz = q.z * q.z, sqw = q.w * q.w;
	const double xy = q.x * q.y, xz = q.x * q.z, xw = q.x * q.w, yz = q.y * q.z, yw = q.y * q.w, zw = q.z * q.w;
	Nuc n;
	n.a1 = {sqx - sqy - sqz + sqw, 2. * (xy + zw), 2. * (xz - yw)};
	n.a2 = {2. * (xy - zw), -sqx + sqy - sqz + sqw, 2. * (yz + xw)};
	n.a3 = {2. * (xz + yw), 2. * (yz - xw), -sqx - sqy + sqz + sqw};
	n.btype = btype;
	n.type = type_of(btype);
	return n;
}

// site offsets from the centre, lab frame
__device__ __forceinline__ dv3 s_back(const Nuc &n) { return n.a1 * (double) POS_BACK_A1 + n.a3 * (double) POS_BACK_A3; }
__device__ __forceinline__ dv3 s_stack(const Nuc &n) { return n.a1 * (double) POS_STACK; }
__device__ __forceinline__ dv3 s_base(const Nuc &n) { return n.a1 * (double) POS_BASE; }
__device__ __forceinline__ dv3 s_stack3(const Nuc &n) { return n.a1 * (double) POS_STACK3_A1 + n.a2 * (double) POS_STACK3_A2; }
__device__ __forceinline__ dv3 s_stack5(const Nuc &n) { return n.a1 * (double) POS_STACK5_A1 + n.a2 * (double) POS_STACK5_A2; }
__device__ __forceinline__ dv3 s_bbv3(const Nuc &n) { return n.a1 * (double) P3_X + n.a2 * (double) P3_Y + n.a3 * (double) P3_Z; }
__device__ __forceinline__ dv3 s_bbv5(const Nuc &n) { return n.a1 * (double) P5_X + n.a2 * (double) P5_Y + n.a3 * (double) P5_Z; }

// (r + qs) - ps
__device__ __forceinline__ dv3 sep(const dv3 &r, const dv3 &qs, const dv3 &ps) { return {(r.x + qs.x) - ps.x, (r.y + qs.y) - ps.y, (r.z + qs.z) - ps.z}; }

__device__ __forceinline__ dv3 min_image(double4 a, double4 b, const oxh_sys &S) {
	double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
	dx = fma(-S.box[0], rint(dx * S.inv_box[0]), dx);
	dy = fma(-S.box[1], rint(dy * S.inv_box[1]), dy);
	dz = fma(-S.box[2], rint(dz * S.inv_box[2]), dz);
	return {dx, dy, dz};
}

// ------------------------------------------------------------ bonded
// FENE, bonded excluded volume and stacking of the bonded pair p -> n3 = q,
// r = q - p.  Force on q into fq (p takes -fq), lab torques on p and q,
// energies into e[0..2]; err bit 0 when the FENE bond is beyond its range.
template<typename A>
__device__ void bonded(const oxh_sys &S, const Nuc &p, const Nuc &q, const dv3 &r, dv3 &fq, dv3 &tp, dv3 &tq, double e[3], int &err) {
	const dv3 pback = s_back(p), qback = s_back(q);
	// backbone
	const dv3 rback = sep(r, qback, pback);
	const double rbackmod = sqrt(dot(rback, rback));
	const double rbackr0 = rbackmod - S.fene_r0_d;
	dv3 force;
	if(S.use_mbf && fabs(rbackr0) > S.mbf_xmax_d) {
		const double xmax = S.mbf_xmax_d, fmax = S.mbf_fmax_d, finf = S.mbf_finf_d;
		const double fene_xmax = -(double) (FENE_EPS / 2.f) * log(1. - SQR(xmax) / (double) FENE_DELTA2);
		const double long_xmax = (fmax - finf) * xmax * log(xmax) + finf * xmax;
		e[0] = (fmax - finf) * xmax * log(fabs(rbackr0)) + finf * fabs(rbackr0) - long_xmax + fene_xmax;
		force = rback * (-copysign(1., rbackr0) * ((fmax - finf) * xmax / fabs(rbackr0) + finf) / rbackmod);
	}
	else {
		double x = rbackr0;
		if(fabs(x) > (double) FENE_DELTA - DBL_EPSILON) {
			err |= 1;
			x = copysign(0.999 * (double) FENE_DELTA, x);
		}
		e[0] = -(double) FENE_EPS * 0.5 * log(1 - SQR(x) / (double) FENE_DELTA2);
		force = rback * (-((double) FENE_EPS * x / ((double) FENE_DELTA2 - SQR(x))) / rbackmod);
	}
	fq = force;
	tp = -cross(pback, force);
	tq = cross(qback, force);

	// bonded excluded volume: base-base, base(p)-back(q), back(p)-base(q)
	const dv3 pbase = s_base(p), qbase = s_base(q);
	dv3 fx;
	double ex = repulsive_lj(sep(r, qbase, pbase), fx, EXCL_S2, EXCL_R2, EXCL_B2, EXCL_RC2);
	fq += fx;
	tp -= cross(pbase, fx);
	tq += cross(qbase, fx);
	ex += repulsive_lj(sep(r, qback, pbase), fx, EXCL_S3, EXCL_R3, EXCL_B3, EXCL_RC3);
	fq += fx;
	tp -= cross(pbase, fx);
	tq += cross(qback, fx);
	ex += repulsive_lj(sep(r, qbase, pback), fx, EXCL_S4, EXCL_R4, EXCL_B4, EXCL_RC4);
	fq += fx;
	tp -= cross(pback, fx);
	tq += cross(qbase, fx);
	e[1] = ex;

	// stacking between STACK_3 of p and STACK_5 of q, with the backbone
	// angles tB1, tB2 and the phi factors on the backbone-backbone vector
	const dv3 pst = s_stack3(p), qst = s_stack5(q);
	const dv3 rstack = sep(r, qst, pst);
	const double rstackmod = sqrt(dot(rstack, rstack));
	const dv3 rstackdir = rstack * (1. / rstackmod);
	const dv3 rbdir = rback * (1. / rbackmod);
	const double cost5 = dot(p.a3, rstackdir);
	const double cost6 = -dot(q.a3, rstackdir);
	const A t5 = lracos<A>(cost5), t6 = lracos<A>(cost6);
	const dv3 pb3 = s_bbv3(p), qb5 = s_bbv5(q);
	const double costB1 = -dot(rbdir, pb3);
	const double costB2 = -dot(rbdir, qb5);
	const double cosphi1 = dot(p.a2, rback) / rbackmod;
	const double cosphi2 = dot(q.a2, rback) / rbackmod;
	const double ff1 = f1<1>(S, rstackmod, q.type, p.type);
	const A pt5 = (A) PIf - t5;
	const double f4t5 = f4<STCK_T5, A>(pt5);
	const double f4t6 = f4<STCK_T6, A>(t6);
	const A tB1 = lracos<A>(costB1), tB2 = lracos<A>(costB2);
	const double f4tB1 = f4<STCK_TB1, A>(tB1);
	const double f4tB2 = f4<STCK_TB2, A>(tB2);
	const double f5phi1 = f5<STCK_PHI1>(cosphi1);
	const double f5phi2 = f5<STCK_PHI2>(cosphi2);
	const double es = ff1 * f4t5 * f4t6 * f5phi1 * f5phi2 * f4tB1 * f4tB2;
	e[2] = es;
	if(es != 0.) {
		const double ff1D = f1D<1>(S, rstackmod, q.type, p.type);
		const double f4t5Dsin = f4Dsin<STCK_T5, A>(pt5);
		const double f4t6Dsin = f4Dsin<STCK_T6, A>(t6);
		const double f4tB1Dsin = f4Dsin<STCK_TB1, A>(tB1);
		const double f4tB2Dsin = f4Dsin<STCK_TB2, A>(tB2);
		const double f5phi1D = f5D<STCK_PHI1>(cosphi1);
		const double f5phi2D = f5D<STCK_PHI2>(cosphi2);

		dv3 force = rstackdir * (-(ff1D * f4t5 * f4t6 * f5phi1 * f5phi2 * f4tB1 * f4tB2));
		force += (p.a3 - rstackdir * cost5) * (-1. / rstackmod) * (ff1 * f4t5Dsin * f4t6 * f5phi1 * f5phi2 * f4tB1 * f4tB2);
		force += (q.a3 + rstackdir * cost6) * (-1. / rstackmod) * (ff1 * f4t5 * f4t6Dsin * f5phi1 * f5phi2 * f4tB1 * f4tB2);
		dv3 pbf = (pb3 + rbdir * costB1) * (-(ff1 * f4t5 * f4t6 * f5phi1 * f5phi2 * f4tB1Dsin * f4tB2 / rbackmod));
		pbf += (qb5 + rbdir * costB2) * (-(ff1 * f4t5 * f4t6 * f5phi1 * f5phi2 * f4tB1 * f4tB2Dsin / rbackmod));
		const double fpp1 = -ff1 * f4t5 * f4t6 * f5phi1D * f5phi2 * f4tB1 * f4tB2;
		pbf += (p.a2 - rback * (cosphi1 / rbackmod)) * (fpp1 / rbackmod);
		// the reference's phi2 factor leaves out f4(tB1) f4(tB2)
		const double fpp2 = -ff1 * f4t5 * f4t6 * f5phi1 * f5phi2D;
		pbf += (q.a2 - rback * (cosphi2 / rbackmod)) * (fpp2 / rbackmod);

		fq += force;
		fq += pbf;
		tp -= cross(pst, force);
		tq += cross(qst, force);
		tp -= cross(pback, pbf);
		tq += cross(qback, pbf);
		tp -= cross(rstackdir, p.a3) * (-ff1 * f4t5Dsin * f4t6 * f5phi1 * f5phi2 * f4tB1 * f4tB2);
		tq += cross(rstackdir, q.a3) * (ff1 * f4t5 * f4t6Dsin * f5phi1 * f5phi2 * f4tB1 * f4tB2);
		tp += cross(rbdir, pb3) * (ff1 * f4t5 * f4t6 * f5phi1 * f5phi2 * f4tB1Dsin * f4tB2);
		tq += cross(rbdir, qb5) * (ff1 * f4t5 * f4t6 * f5phi1 * f5phi2 * f4tB1 * f4tB2Dsin);
		tp += cross(p.a2, rbdir) * fpp1;
		tq += cross(q.a2, rbdir) * fpp2;
	}
}

// ------------------------------------------------------- non-bonded
// Excluded volume, hydrogen bonding, cross stacking and coaxial stacking of
// the non-bonded ordered pair (p, q), r = q - p.  Adds the force on q to fq
// and the lab torque on p to tp; energies into e[0..3].
template<typename A>
__device__ void nonbonded(const oxh_sys &S, const Nuc &p, const Nuc &q, const dv3 &r, dv3 &fq, dv3 &tp, double e[4]) {
	const dv3 pback = s_back(p), qback = s_back(q), pbase = s_base(p), qbase = s_base(q);
	// excluded volume of four site pairs
	dv3 fx;
	const dv3 rbb = sep(r, qbase, pbase);
	double ex = repulsive_lj(rbb, fx, EXCL_S2, EXCL_R2, EXCL_B2, EXCL_RC2);
	fq += fx;
	tp -= cross(pbase, fx);
	ex += repulsive_lj(sep(r, qback, pbase), fx, EXCL_S3, EXCL_R3, EXCL_B3, EXCL_RC3);
	fq += fx;
	tp -= cross(pbase, fx);
	ex += repulsive_lj(sep(r, qbase, pback), fx, EXCL_S4, EXCL_R4, EXCL_B4, EXCL_RC4);
	fq += fx;
	tp -= cross(pback, fx);
	const dv3 rback = sep(r, qback, pback);
	ex += repulsive_lj(rback, fx, EXCL_S1, EXCL_R1, EXCL_B1, EXCL_RC1);
	fq += fx;
	tp -= cross(pback, fx);
	e[0] = ex;
	e[1] = e[2] = e[3] = 0.;

	const dv3 &a1 = p.a1, &a3 = p.a3, &b1 = q.a1, &b3 = q.a3;
	const double rbmod2 = dot(rbb, rbb);
	// hydrogen bonding (Watson-Crick-like pairs: base types summing to 3)
	if(q.btype + p.btype == 3 && SQR((double) HYDR_RCLOW) < rbmod2 && rbmod2 < SQR((double) HYDR_RCHIGH)) {
		const double rhmod = sqrt(rbmod2);
		if((double) HYDR_RCLOW < rhmod && rhmod < (double) HYDR_RCHIGH) {
			const dv3 rhdir = rbb * (1. / rhmod);
			const double cost1 = -dot(a1, b1), cost2 = -dot(b1, rhdir), cost3 = dot(a1, rhdir);
			const double cost4 = dot(a3, b3), cost7 = -dot(b3, rhdir), cost8 = dot(a3, rhdir);
			const A t1 = lracos<A>(cost1), t2 = lracos<A>(cost2), t3 = lracos<A>(cost3);
			const A t4 = lracos<A>(cost4), t7 = lracos<A>(cost7), t8 = lracos<A>(cost8);
			const double ff1 = f1<0>(S, rhmod, q.type, p.type);
			const double f4t1 = f4<HYDR_T1, A>(t1), f4t2 = f4<HYDR_T2, A>(t2), f4t3 = f4<HYDR_T3, A>(t3);
			const double f4t4 = f4<HYDR_T4, A>(t4), f4t7 = f4<HYDR_T7, A>(t7), f4t8 = f4<HYDR_T8, A>(t8);
			const double eh = ff1 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8;
			e[1] = eh;
			if(eh != 0.) {
				const double ff1D = f1D<0>(S, rhmod, q.type, p.type);
				// d f4(acos c) / dc = -f4Dsin, with the signs the reference's
				// mesh derivatives carry
				const double f4t1Dsin = -f4Dsin<HYDR_T1, A>(t1);
				const double f4t2Dsin = -f4Dsin<HYDR_T2, A>(t2);
				const double f4t3Dsin = f4Dsin<HYDR_T3, A>(t3);
				const double f4t4Dsin = f4Dsin<HYDR_T4, A>(t4);
				const double f4t7Dsin = -f4Dsin<HYDR_T7, A>(t7);
				const double f4t8Dsin = f4Dsin<HYDR_T8, A>(t8);
				dv3 force = rhdir * (-ff1D * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8);
				dv3 tl = cross(a3, b3) * (-(-ff1 * f4t1 * f4t2 * f4t3 * f4t4Dsin * f4t7 * f4t8));
				tl -= cross(a1, b1) * (-ff1 * f4t1Dsin * f4t2 * f4t3 * f4t4 * f4t7 * f4t8);
				force += (b1 + rhdir * cost2) * (1. / rhmod) * (ff1 * f4t1 * f4t2Dsin * f4t3 * f4t4 * f4t7 * f4t8);
				force += (a1 - rhdir * cost3) * (1. / rhmod) * (ff1 * f4t1 * f4t2 * f4t3Dsin * f4t4 * f4t7 * f4t8);
				tl += cross(rhdir, a1) * (-ff1 * f4t1 * f4t2 * f4t3Dsin * f4t4 * f4t7 * f4t8);
				force += (b3 + rhdir * cost7) * (1. / rhmod) * (ff1 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7Dsin * f4t8);
				force += (a3 - rhdir * cost8) * (1. / rhmod) * (ff1 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8Dsin);
				tl += cross(rhdir, a3) * (-ff1 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8Dsin);
				fq += force;
				tp += tl;
				tp -= cross(pbase, force);
			}
		}
	}

	// cross stacking, on the base-base separation
	if(SQR((double) F2_RCLOW[0]) < rbmod2 && rbmod2 < SQR((double) F2_RCHIGH[0])) {
		const double rcmod = sqrt(rbmod2);
		if((double) F2_RCLOW[0] < rcmod && rcmod < (double) F2_RCHIGH[0]) {
			const dv3 rcdir = rbb * (1. / rcmod);
			const double cost2 = -dot(b1, rcdir), cost3 = dot(a1, rcdir), cost7 = -dot(b3, rcdir), cost8 = dot(a3, rcdir);
			const A t1 = lracos<A>(-dot(a1, b1)), t2 = lracos<A>(cost2), t3 = lracos<A>(cost3);
			const A t7 = lracos<A>(cost7), t8 = lracos<A>(cost8);
			const double ff2 = f2<0>(rcmod);
			const double f4t1 = f4<CRST_T1, A>(t1), f4t2 = f4<CRST_T2, A>(t2), f4t3 = f4<CRST_T3, A>(t3);
			const double f4t7 = f4<CRST_T7, A>(t7) + f4<CRST_T7, A>((A) PIf - t7);
			const double f4t8 = f4<CRST_T8, A>(t8) + f4<CRST_T8, A>((A) PIf - t8);
			const double ec = ff2 * f4t1 * f4t2 * f4t3 * f4t7 * f4t8;
			e[2] = ec;
			if(ec != 0.) {
				const double ff2D = f2D<0>(rcmod);
				const double f4t1Dsin = -f4Dsin<CRST_T1, A>(t1);
				const double f4t2Dsin = -f4Dsin<CRST_T2, A>(t2);
				const double f4t3Dsin = f4Dsin<CRST_T3, A>(t3);
				const double f4t7Dsin = -f4Dsin<CRST_T7, A>(t7) + f4Dsin<CRST_T7, A>((A) PIf - t7);
				const double f4t8Dsin = f4Dsin<CRST_T8, A>(t8) - f4Dsin<CRST_T8, A>((A) PIf - t8);
				// the reference's theta4 factor is 1 here, its derivative 0
				dv3 force = rcdir * (-(ff2D * f4t1 * f4t2 * f4t3 * f4t7 * f4t8));
				dv3 tl = cross(a1, b1) * (-(-ff2 * f4t1Dsin * f4t2 * f4t3 * f4t7 * f4t8));
				force += (b1 + rcdir * cost2) * (ff2 * f4t1 * f4t2Dsin * f4t3 * f4t7 * f4t8 / rcmod);
				force += (a1 - rcdir * cost3) * (ff2 * f4t1 * f4t2 * f4t3Dsin * f4t7 * f4t8 / rcmod);
				tl += cross(rcdir, a1) * (-ff2 * f4t1 * f4t2 * f4t3Dsin * f4t7 * f4t8);
				force += (b3 + rcdir * cost7) * (1. / rcmod) * (ff2 * f4t1 * f4t2 * f4t3 * f4t7Dsin * f4t8);
				force += (a3 - rcdir * cost8) * (1. / rcmod) * (ff2 * f4t1 * f4t2 * f4t3 * f4t7 * f4t8Dsin);
				tl += cross(rcdir, a3) * (-ff2 * f4t1 * f4t2 * f4t3 * f4t7 * f4t8Dsin);
				fq += force;
				tp += tl;
				tp -= cross(pbase, force);
			}
		}
	}

	// coaxial stacking, on the stack-stack separation
	const dv3 pst = s_stack(p), qst = s_stack(q);
	const dv3 rstack = sep(r, qst, pst);
	const double rs2 = dot(rstack, rstack);
	if(SQR((double) F2_RCLOW[1]) < rs2 && rs2 < SQR((double) F2_RCHIGH[1])) {
		const double rsmod = sqrt(rs2);
		if((double) F2_RCLOW[1] < rsmod && rsmod < (double) F2_RCHIGH[1]) {
			const dv3 rsdir = rstack * (1. / rsmod);
			const double cost5 = dot(a3, rsdir), cost6 = -dot(b3, rsdir);
			const A t1 = lracos<A>(-dot(a1, b1)), t4 = lracos<A>(dot(a3, b3)), t5 = lracos<A>(cost5), t6 = lracos<A>(cost6);
			const double rbmod = sqrt(dot(rback, rback));
			const dv3 rbdir = rback * (1. / rbmod);
			const double cosphi3 = dot(rsdir, cross(rbdir, a1));
			const double cosphi4 = dot(rsdir, cross(rbdir, b1));
			const double ff2 = f2<1>(rsmod);
			const A t1b = (A) (2 * PIf) - t1;
			const double f4t1 = f4<CXST_T1, A>(t1) + f4<CXST_T1, A>(t1b);
			const double f4t4 = f4<CXST_T4, A>(t4);
			const double f4t5 = f4<CXST_T5, A>(t5) + f4<CXST_T5, A>((A) PIf - t5);
			const double f4t6 = f4<CXST_T6, A>(t6) + f4<CXST_T6, A>((A) PIf - t6);
			const double f5c3 = f5<CXST_PHI3>(cosphi3);
			const double f5c4 = f5<CXST_PHI4>(cosphi4);
			const double ecx = ff2 * f4t1 * f4t4 * f4t5 * f4t6 * f5c3 * f5c4;
			e[3] = ecx;
			if(ecx != 0.) {
				const double ff2D = f2D<1>(rsmod);
				const double f4t1Dsin = -f4Dsin<CXST_T1, A>(t1) + f4Dsin<CXST_T1, A>(t1b);
				const double f4t4Dsin = f4Dsin<CXST_T4, A>(t4);
				const double f4t5Dsin = f4Dsin<CXST_T5, A>(t5) - f4Dsin<CXST_T5, A>((A) PIf - t5);
				const double f4t6Dsin = -f4Dsin<CXST_T6, A>(t6) + f4Dsin<CXST_T6, A>((A) PIf - t6);
				const double f5Dc3 = f5D<CXST_PHI3>(cosphi3);
				const double f5Dc4 = f5D<CXST_PHI4>(cosphi4);
				dv3 force = rsdir * (-ff2D * f4t1 * f4t4 * f4t5 * f4t6 * f5c3 * f5c4);
				dv3 tl = cross(a1, b1) * (-(-ff2 * f4t1Dsin * f4t4 * f4t5 * f4t6 * f5c3 * f5c4));
				tl -= cross(a3, b3) * (-ff2 * f4t1 * f4t4Dsin * f4t5 * f4t6 * f5c3 * f5c4);
				const double fact5 = ff2 * f4t1 * f4t4 * f4t5Dsin * f4t6 * f5c3 * f5c4;
				force += (a3 - rsdir * cost5) * (fact5 / rsmod);
				tl -= cross(rsdir, a3) * fact5;
				const double fact6 = ff2 * f4t1 * f4t4 * f4t5 * f4t6Dsin * f5c3 * f5c4;
				force += (b3 + rsdir * cost6) * (1. / rsmod) * fact6;
				tl -= cross(pst, force);

				// phi3 (on a1) and phi4 (on b1): forces on the stack and
				// backbone sites, torques on p as the reference collects them
				double force_c = ff2 * f4t1 * f4t4 * f4t5 * f4t6 * f5c4 * f5Dc3;
				dv3 rba1 = cross(rbdir, a1);
				dv3 a1rs = cross(a1, rsdir);
				double rb_dot_a1rs = dot(rbdir, a1rs);
				double rs_dot_rba1 = dot(rsdir, rba1);
				dv3 fst = (rba1 - rsdir * rs_dot_rba1) * (-(force_c / rsmod));
				dv3 fbk = (a1rs - rbdir * rb_dot_a1rs) * (-(force_c / rbmod));
				dv3 myforce = fst;
				myforce += fbk;
				dv3 mt1 = -cross(pst, fst);
				mt1 -= cross(pback, fbk);
				mt1 -= cross(a1, cross(rsdir, rbdir)) * force_c;

				force_c = ff2 * f4t1 * f4t4 * f4t5 * f4t6 * f5c3 * f5Dc4;
				rba1 = cross(rbdir, b1);
				a1rs = cross(b1, rsdir);
				rb_dot_a1rs = dot(rbdir, a1rs);
				rs_dot_rba1 = dot(rsdir, rba1);
				fst = (rba1 - rsdir * rs_dot_rba1) * (-(force_c / rsmod));
				fbk = (a1rs - rbdir * rb_dot_a1rs) * (-(force_c / rbmod));
				myforce += fst;
				myforce += fbk;
				mt1 -= cross(pst, fst);
				mt1 -= cross(pback, fbk);

				force += myforce;
				tl += mt1;
				fq += force;
				tp += tl;
			}
		}
	}
}

// oxRNA2 Debye-Hueckel between backbone sites; cf = product of the two
// charge factors.  Adds the force on q to fq, returns the energy.
__device__ __forceinline__ double debye_huckel(const oxh_sys &S, const dv3 &rback, double cf, dv3 &fq) {
	const double rbackmod = sqrt(dot(rback, rback));
	double energy = 0.;
	if(rbackmod < S.dh_rc_d) {
		if(rbackmod < S.dh_rhigh_d) energy = exp(rbackmod * S.dh_minus_kappa_d) * (S.dh_prefactor_d / rbackmod);
		else energy = S.dh_b_d * SQR(rbackmod - S.dh_rc_d);
		energy *= cf;
		if(energy != 0.) {
			const dv3 rbdir = rback * (1. / rbackmod);
			dv3 force;
			if(rbackmod < S.dh_rhigh_d)
				force = rbdir * (-1.0 * (S.dh_prefactor_d * exp(S.dh_minus_kappa_d * rbackmod)) * (S.dh_minus_kappa_d / rbackmod - 1.0 / SQR(rbackmod)));
			else force = rbdir * (-(2.0 * S.dh_b_d * (rbackmod - S.dh_rc_d)));
			fq += force * cf;
		}
	}
	return energy;
}

__device__ __forceinline__ void add_virial(double sv[6], const dv3 &r, const dv3 &f) {
	sv[0] += r.x * f.x;
	sv[1] += r.y * f.y;
	sv[2] += r.z * f.z;
	sv[3] += r.x * f.y;
	sv[4] += r.x * f.z;
	sv[5] += r.y * f.z;
}

__device__ __forceinline__ double group_sum(double v, int K) {
	for(int off = 1; off < K; off <<= 1) v += __shfl_xor(v, off);
	return v;
}

__device__ __forceinline__ float4 out4(const dv3 &v, double w, float4 *) { return make_float4((float) v.x, (float) v.y, (float) v.z, (float) w); }
__device__ __forceinline__ double4 out4(const dv3 &v, double w, double4 *) { return make_double4(v.x, v.y, v.z, w); }

// One nucleotide per K lanes.  F: lab force, .w = half of the pair
// energies the nucleotide takes part in; T: body-frame torque, .w = half
// of its hydrogen-bonding energy.  With SPLIT the eight energy terms and
// the virial go to esplit / stress.
template<int K, bool SPLIT, typename OUT4>
__global__ void __launch_bounds__(64) k_forces_rna(int N, const double4 *__restrict__ pos, const double4 *__restrict__ q64,
		const float4 *__restrict__ bpos, const int2 *__restrict__ bonds, const int2 *__restrict__ pinfo, const int *__restrict__ nbr,
		const int2 *__restrict__ nnbr, int max_neigh, const oxh_sys *__restrict__ systems, oxh_sys_state *__restrict__ sstate,
		OUT4 *__restrict__ Fout, OUT4 *__restrict__ Tout, double *__restrict__ esplit, int *__restrict__ hbcount, int kstep,
		const int *__restrict__ limit, int *__restrict__ any, double *__restrict__ stress, int xcd) {
	// angles (acos, sin and the f4 factors) in FP32 in the mixed form
	using A = typename std::conditional<std::is_same<OUT4, float4>::value, float, double>::type;
	const int blk = xcd ? xcd_block(((unsigned) N * K + blockDim.x - 1) / blockDim.x) : (int) blockIdx.x;
	if(blk < 0) return;
	const int t = blk * blockDim.x + threadIdx.x;
	const int i = t / K;
	const int l = t % K;
	if(kstep >= 0 && t == 0) *any = 0;
	if(i >= N) return;
	const int2 pi = pinfo[i];
	if(kstep >= 0 && l == 0 && i == systems[pi.y].first && sstate[pi.y].need_rebuild) {
		sstate[pi.y].need_rebuild = 0;
		sstate[pi.y].n_rebuilds += 1;
	}
	const oxh_sys &S = systems[pi.y];
	if(idle(i, S, pi.y, kstep, limit)) return;
	const double4 P = pos[i];
	const Nuc np = nuc(q64[i], pi.x);
	const int2 bd = bonds[i];

	dv3 F = {0, 0, 0}, T = {0, 0, 0};
	double E = 0., Ehb = 0.;
	double es[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	double sv[6] = {0, 0, 0, 0, 0, 0};
	int err = 0, nhb = 0;

	// bonded neighbours: i is p of the bond i -> n3 (lane 0) and q of the
	// bond n5 -> i (lane 1; with one lane, lane 0 takes both in turn)
	for(int pass = 0; pass < (K > 1 ? 1 : 2); pass++) {
		const bool n3side = K > 1 ? l == 0 : pass == 0;
		const int j = n3side ? bd.x : bd.y;
		if((K > 1 && l > 1) || j < 0) continue;
		const Nuc nj = nuc(q64[j], pinfo[j].x);
		const double4 Pj = pos[j];
		const dv3 r = n3side ? min_image(P, Pj, S) : min_image(Pj, P, S);
		dv3 fq = {0, 0, 0}, tp = {0, 0, 0}, tq = {0, 0, 0};
		double e3[3];
		bonded<A>(S, n3side ? np : nj, n3side ? nj : np, r, fq, tp, tq, e3, err);
		if(n3side) {
			F -= fq;
			T += tp;
		}
		else {
			F += fq;
			T += tq;
		}
		E += 0.5 * (e3[0] + e3[1] + e3[2]);
		if(SPLIT) {
			es[0] += 0.5 * e3[0];
			es[1] += 0.5 * e3[1];
			es[2] += 0.5 * e3[2];
			add_virial(sv, r, fq);
		}
	}

	const int2 nn = nnbr[i];
	const int *row = nbr + (size_t) i * max_neigh;
	const double cf_i = bpos[i].w;
	const dv3 pback = s_back(np);
	for(int k = l; k < nn.x + nn.y; k += K) {
		const int j = k < nn.x ? row[k] : row[max_neigh - 1 - (k - nn.x)];
		const Nuc nj = nuc(q64[j], pinfo[j].x);
		const dv3 r = min_image(P, pos[j], S);
		dv3 fq = {0, 0, 0}, tp = {0, 0, 0};
		double e = 0.;
		if(k < nn.x) {
			double e4[4];
			nonbonded<A>(S, np, nj, r, fq, tp, e4);
			e = e4[0] + e4[1] + e4[2] + e4[3];
			Ehb += 0.5 * e4[1];
			if(hbcount && e4[1] < -0.1) nhb++;
			if(SPLIT) {
				es[3] += 0.5 * e4[0];
				es[4] += 0.5 * e4[1];
				es[5] += 0.5 * e4[2];
				es[6] += 0.5 * e4[3];
			}
		}
		if(S.dh_rc_d > 0.) {
			dv3 fdh = {0, 0, 0};
			const double edh = debye_huckel(S, sep(r, s_back(nj), pback), cf_i * (double) bpos[j].w, fdh);
			fq += fdh;
			tp -= cross(pback, fdh);
			e += edh;
			if(SPLIT) es[7] += 0.5 * edh;
		}
		F -= fq;
		T += tp;
		E += 0.5 * e;
		if(SPLIT) add_virial(sv, r, fq);
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
		if(SPLIT) {
			for(int q = 0; q < 8; q++) es[q] = group_sum(es[q], K);
			for(int q = 0; q < 6; q++) sv[q] = group_sum(sv[q], K);
		}
		if(hbcount) nhb = (int) group_sum((double) nhb, K);
		err = (int) group_sum((double) err, K);
	}
	if(l != 0) return;
	const dv3 Tb = {dot(np.a1, T), dot(np.a2, T), dot(np.a3, T)};
	Fout[i] = out4(F, E, Fout);
	Tout[i] = out4(Tb, Ehb, Tout);
	if(SPLIT) {
		for(int q = 0; q < 8; q++) esplit[(size_t) q * N + i] = es[q];
		for(int q = 0; q < 6; q++) stress[(size_t) q * N + i] = 0.5 * sv[q];
	}
	if(hbcount) hbcount[i] = nhb;
	if(err) atomicOr(&sstate[pi.y].error, ERR_BOND);
}

template<int K, typename OUT4>
void launch_k(hipStream_t st, const Batch &b, OUT4 *F, OUT4 *T, bool split, int k) {
	const int block = 64;
	const unsigned nblk = (unsigned) (((size_t) b.N * K + block - 1) / block);
	dim3 grid(b.xcd ? xcd_grid(nblk) : nblk);
	const int *lim = k >= 0 ? b.limit : nullptr;
	if(split)
		k_forces_rna<K, true, OUT4><<<grid, block, 0, st>>>(b.N, b.pos, b.q64, b.bpos, b.bonds, b.pinfo, b.nbr, b.nnbr, b.max_neigh, b.d_sys,
				b.d_state, F, T, b.esplit, b.hbcount, k, lim, b.any_rebuild, b.stress, b.xcd);
	else
		k_forces_rna<K, false, OUT4><<<grid, block, 0, st>>>(b.N, b.pos, b.q64, b.bpos, b.bonds, b.pinfo, b.nbr, b.nnbr, b.max_neigh, b.d_sys,
				b.d_state, F, T, nullptr, b.want_hbcount ? b.hbcount : nullptr, k, lim, b.any_rebuild, nullptr, b.xcd);
}

template<typename OUT4>
void launch_lanes(hipStream_t st, const Batch &b, OUT4 *F, OUT4 *T, bool split, int k) {
	switch(b.lanes) {
	case 1: launch_k<1>(st, b, F, T, split, k); break;
	case 2: launch_k<2>(st, b, F, T, split, k); break;
	case 8: launch_k<8>(st, b, F, T, split, k); break;
	case 16: launch_k<16>(st, b, F, T, split, k); break;
	default: launch_k<4>(st, b, F, T, split, k); break;
	}
}

// FFS order-parameter quantities of listed oxRNA pairs (p, q, same
// strand) of system s, in the system's upload order: the hydrogen-bonding
// energy (zero for bonded neighbours), or the base-base / centre-centre
// distance (the plain difference for pairs of one strand, the nearest image
// otherwise), as src/Utilities/OrderParameters.cpp:73-98 takes them
__global__ void k_rna_pair_quantities(int n, int kind, int s, const int4 *__restrict__ pairs, const double4 *__restrict__ pos,
		const double4 *__restrict__ q64, const int2 *__restrict__ pinfo, const int2 *__restrict__ bonds, const int *__restrict__ slot_of,
		const oxh_sys *__restrict__ systems, double *__restrict__ out) {
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if(k >= n) return;
	const oxh_sys &S = systems[s];
	const int4 pr = pairs[k];
	const int p = slot_of[S.first + pr.x], q = slot_of[S.first + pr.y];
	const double4 P = pos[p], Q = pos[q];
	const bool plain = kind >= OXH_PAIR_BASE_DISTANCE && pr.z != 0;
	const dv3 r = plain ? dv3{Q.x - P.x, Q.y - P.y, Q.z - P.z} : min_image(P, Q, S);
	if(kind == OXH_PAIR_CENTRE_DISTANCE) {
		out[k] = sqrt(dot(r, r));
		return;
	}
	const Nuc np = nuc(q64[p], pinfo[p].x), nq = nuc(q64[q], pinfo[q].x);
	if(kind == OXH_PAIR_BASE_DISTANCE) {
		const dv3 rb = sep(r, s_base(nq), s_base(np));
		out[k] = sqrt(dot(rb, rb));
		return;
	}
	const int2 bd = bonds[p];
	double e4[4] = {0., 0., 0., 0.};
	if(p != q && bd.x != q && bd.y != q) {
		dv3 fq = {0, 0, 0}, tp = {0, 0, 0};
		nonbonded<double>(S, np, nq, r, fq, tp, e4);
	}
	out[k] = e4[1];
}

}  // namespace rna

void launch_rna_pair_quantities(hipStream_t st, const Batch &b, int sys, int kind, int n, const int4 *d_pairs, double *d_out) {
	if(n <= 0) return;
	rna::k_rna_pair_quantities<<<(n + 255) / 256, 256, 0, st>>>(n, kind, sys, d_pairs, b.pos, b.q64, b.pinfo, b.bonds, b.slot_of, b.d_sys,
			d_out);
}

void launch_rna_forces(hipStream_t st, const Batch &b, bool split, int k) {
	if(b.N == 0) return;
	if(b.fp64) rna::launch_lanes(st, b, b.F64, b.T64, split, k);
	else rna::launch_lanes(st, b, b.F, b.T, split, k);
}

}  // namespace oxh
