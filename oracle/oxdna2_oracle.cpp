/*
 * FP64 CPU oracle for the oxDNA2 MD hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker.  The product (oxdna_amd/) never links, loads or calls it.
 *
 * It restates, in plain FP64 C++, the algorithm of the reference's CPU
 * backend with `interaction_type = DNA2_nomesh` (analytic f4), which is the
 * reference's own FP64 counterpart of its GPU path
 * (src/Interactions/InteractionFactory.cpp:50-57):
 *   - model constants and their T / salt dependence
 *       DNAInteraction::init            src/Interactions/DNAInteraction.cpp:298-377
 *       DNA2Interaction::get_settings   src/Interactions/DNA2Interaction.cpp:62-101
 *       DNA2Interaction::init           src/Interactions/DNA2Interaction.cpp:103-155
 *   - radial / angular modulation      src/Interactions/DNAInteraction.cpp:1249-1454,
 *                                      src/Interactions/DNA2Interaction.cpp:308-363
 *   - pair terms                       src/Interactions/DNAInteraction.cpp:415-1205,
 *                                      src/Interactions/DNA2Interaction.cpp:157-306
 *   - the force loop                   src/Backends/MD_CPUBackend.cpp:146-167
 *   - energy split                     src/Interactions/BaseInteraction.cpp:61-86
 * plus the pieces of the GPU path whose semantics are this build's own
 * (FP64 Verlet-pair predicate, Hilbert key, Philox counter RNG, thermostats
 * driven by given noise) so that the HIP kernels can be checked exactly.
 *
 * Pinning (what keeps this restatement honest) lives in tests/: the
 * reference's own fixture test/DNA/FORCE_FIELD/AVG_SEQ/reference.dat, and
 * full-precision energies / forces / torques written by the reference binary
 * built from source (oracle/ref.mk, fixtures in tests/golden/), plus
 * finite-difference, Newton's-third-law and closed-form checks.
 *
 * Conventions: a1, a2, a3 are the rows of the reference's orientationT
 * (principal, third and stacking axes); torques are returned in the body
 * frame (p->torque += p->orientationT * torque_lab), forces in the lab frame.
 */
#include "oxdna2_model.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

using namespace oxmodel;

namespace {

struct V3 {
	double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(double s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, double s) { double f = 1. / s; return {a.x * f, a.y * f, a.z * f}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 &operator+=(V3 &a, V3 b) { a = a + b; return a; }
inline V3 &operator-=(V3 &a, V3 b) { a = a - b; return a; }
inline double SQR(double x) { return x * x; }

}  // namespace

extern "C" {

/* Model parameters of one system.  Filled by oxo_params_init. */
typedef struct {
	double T, salt;
	double F1_EPS[2][5][5], F1_SHIFT[2][5][5];
	double F1_A[2], F1_RC[2], F1_R0[2], F1_BLOW[2], F1_BHIGH[2], F1_RLOW[2], F1_RHIGH[2], F1_RCLOW[2], F1_RCHIGH[2];
	double F2_K[2], F2_RC[2], F2_R0[2], F2_BLOW[2], F2_BHIGH[2], F2_RLOW[2], F2_RHIGH[2], F2_RCLOW[2], F2_RCHIGH[2];
	double dh_prefactor, dh_lambdafactor, dh_RHIGH, dh_RC, dh_B, dh_minus_kappa;
	int dh_half_charged_ends;
	double hb_multiplier;
	int use_mbf;
	double mbf_fmax, mbf_finf, mbf_xmax;
	double fene_r0;
	double rcut, sqr_rcut;
} oxo_params;

}  // extern "C"

namespace {

/* --- modulation functions (src/Interactions/DNAInteraction.cpp:1249-1454) --- */

double f1(const oxo_params *P, double r, int type, int n3, int n5) {
	double val = 0.;
	if(r < P->F1_RCHIGH[type]) {
		if(r > P->F1_RHIGH[type]) {
			val = P->F1_EPS[type][n3][n5] * P->F1_BHIGH[type] * SQR(r - P->F1_RCHIGH[type]);
		}
		else if(r > P->F1_RLOW[type]) {
			double tmp = 1 - std::exp(-(r - P->F1_R0[type]) * P->F1_A[type]);
			val = P->F1_EPS[type][n3][n5] * SQR(tmp) - P->F1_SHIFT[type][n3][n5];
		}
		else if(r > P->F1_RCLOW[type]) {
			val = P->F1_EPS[type][n3][n5] * P->F1_BLOW[type] * SQR(r - P->F1_RCLOW[type]);
		}
	}
	return val;
}

double f1D(const oxo_params *P, double r, int type, int n3, int n5) {
	double val = 0.;
	if(r < P->F1_RCHIGH[type]) {
		if(r > P->F1_RHIGH[type]) {
			val = 2 * P->F1_BHIGH[type] * (r - P->F1_RCHIGH[type]);
		}
		else if(r > P->F1_RLOW[type]) {
			double tmp = std::exp(-(r - P->F1_R0[type]) * P->F1_A[type]);
			val = 2 * (1 - tmp) * tmp * P->F1_A[type];
		}
		else if(r > P->F1_RCLOW[type]) {
			val = 2 * P->F1_BLOW[type] * (r - P->F1_RCLOW[type]);
		}
	}
	return P->F1_EPS[type][n3][n5] * val;
}

double f2(const oxo_params *P, double r, int type) {
	double val = 0.;
	if(r < P->F2_RCHIGH[type]) {
		if(r > P->F2_RHIGH[type]) {
			val = P->F2_K[type] * P->F2_BHIGH[type] * SQR(r - P->F2_RCHIGH[type]);
		}
		else if(r > P->F2_RLOW[type]) {
			val = (P->F2_K[type] / 2.) * (SQR(r - P->F2_R0[type]) - SQR(P->F2_RC[type] - P->F2_R0[type]));
		}
		else if(r > P->F2_RCLOW[type]) {
			val = P->F2_K[type] * P->F2_BLOW[type] * SQR(r - P->F2_RCLOW[type]);
		}
	}
	return val;
}

double f2D(const oxo_params *P, double r, int type) {
	double val = 0.;
	if(r < P->F2_RCHIGH[type]) {
		if(r > P->F2_RHIGH[type]) {
			val = 2. * P->F2_K[type] * P->F2_BHIGH[type] * (r - P->F2_RCHIGH[type]);
		}
		else if(r > P->F2_RLOW[type]) {
			val = P->F2_K[type] * (r - P->F2_R0[type]);
		}
		else if(r > P->F2_RCLOW[type]) {
			val = 2. * P->F2_K[type] * P->F2_BLOW[type] * (r - P->F2_RCLOW[type]);
		}
	}
	return val;
}

double f4_of_theta(double t, int s) {
	double val = 0.;
	t -= (double) F4_T0[s];
	if(t < 0) t *= -1;
	if(t < (double) F4_TC[s]) {
		if(t > (double) F4_TS[s]) val = (double) F4_B[s] * SQR((double) F4_TC[s] - t);
		else val = 1. - (double) F4_A[s] * SQR(t);
	}
	return val;
}

/* (df4/dtheta) / sin(theta), with the reference's small-sine guard
 * (src/Interactions/DNAInteraction.cpp:1394-1420) */
double f4Dsin_of_theta(double t, int s) {
	double val = 0.;
	double m = 1.;
	double tt0 = t - (double) F4_T0[s];
	if(tt0 < 0) {
		tt0 *= -1;
		m = -1.;
	}
	if(tt0 < (double) F4_TC[s]) {
		double sint = std::sin(t);
		if(tt0 > (double) F4_TS[s]) {
			val = m * 2 * (double) F4_B[s] * (tt0 - (double) F4_TC[s]) / sint;
		}
		else {
			if(SQR(sint) > 1e-8) val = -m * 2 * (double) F4_A[s] * tt0 / sint;
			else val = -m * 2 * (double) F4_A[s];
		}
	}
	return val;
}

// oxDNA2 extra term on coaxial theta1 (src/Interactions/DNA2Interaction.cpp:320-351)
double f4_pure_harmonic(double t) {
	t -= (double) CXST_THETA1_SB;
	return (t < 0) ? 0. : (double) CXST_THETA1_SA * SQR(t);
}
double f4Dsin_pure_harmonic(double t) {
	double tt0 = t - (double) CXST_THETA1_SB;
	if(tt0 < 0) return 0.;
	double sint = std::sin(t);
	if(SQR(sint) > 1e-8) return 2 * (double) CXST_THETA1_SA * tt0 / sint;
	return 2 * (double) CXST_THETA1_SA;
}

double safe_acos(double x) { return std::acos(std::fmax(-1.0, std::fmin(1.0, x))); }

/* _custom_f4 / _custom_f4D of DNA2Interaction_nomesh
 * (src/Interactions/DNA2Interaction.h:121-152,
 *  src/Interactions/DNAInteraction.cpp:1297-1315,
 *  src/Interactions/DNA2Interaction.cpp:308-318) */
double custom_f4(double cost, int s) {
	if(s != F4_CXST_T1) return f4_of_theta(safe_acos(cost), s);
	if(cost * cost > 1) cost = std::copysign(1., cost);
	double t = std::acos(cost);
	return f4_of_theta(t, s) + f4_pure_harmonic(t);
}
double custom_f4D(double cost, int s) {
	if(s != F4_CXST_T1) return -f4Dsin_of_theta(safe_acos(cost), s);
	if(cost * cost > 1) cost = std::copysign(1., cost);
	double t = std::acos(cost);
	return -f4Dsin_of_theta(t, s) - f4Dsin_pure_harmonic(t);
}

double f5(double f, int s) {
	double val = 0.;
	if(f > (double) F5_XC[s]) {
		if(f < (double) F5_XS[s]) val = (double) F5_B[s] * SQR((double) F5_XC[s] - f);
		else if(f < 0) val = 1. - (double) F5_A[s] * SQR(f);
		else val = 1.;
	}
	return val;
}
double f5D(double f, int s) {
	double val = 0.;
	if(f > (double) F5_XC[s]) {
		if(f < (double) F5_XS[s]) val = 2 * (double) F5_B[s] * (f - (double) F5_XC[s]);
		else if(f < 0) val = -2. * (double) F5_A[s] * f;
	}
	return val;
}

/* truncated/smoothed LJ (src/Interactions/DNAInteraction.cpp:1182-1205) */
double repulsive_lj(V3 r, V3 &force, double sigma, double rstar, double b, double rc) {
	double rnorm = SQR(r.x) + SQR(r.y) + SQR(r.z);
	double energy = 0.;
	force = {0, 0, 0};
	if(rnorm < SQR(rc)) {
		if(rnorm > SQR(rstar)) {
			double rmod = std::sqrt(rnorm);
			double rrc = rmod - rc;
			energy = (double) EXCL_EPS * b * SQR(rrc);
			force = -r * (2 * (double) EXCL_EPS * b * rrc / rmod);
		}
		else {
			double tmp = SQR(sigma) / rnorm;
			double lj_part = tmp * tmp * tmp;
			energy = 4 * (double) EXCL_EPS * (SQR(lj_part) - lj_part);
			force = -r * (24 * (double) EXCL_EPS * (lj_part - 2 * SQR(lj_part)) / rnorm);
		}
	}
	if(energy == 0.) force = {0, 0, 0};
	return energy;
}

/* One nucleotide as the pair terms see it. */
struct Nuc {
	V3 pos, a1, a2, a3;
	V3 back, stack, base;  // interaction centres relative to pos
	int btype, type;
	int n3, n5;
	int index;
};

/* DNANucleotide::set_positions for DNA2 (grooving = 1),
 * src/Particles/DNANucleotide.cpp:66-88 */
void set_sites(Nuc &p) {
	p.back = p.a1 * (double) POS_MM_BACK1 + p.a2 * (double) POS_MM_BACK2;
	p.stack = p.a1 * (double) POS_STACK;
	p.base = p.stack * ((double) POS_BASE / (double) POS_STACK);
}

/* Accumulators: lab-frame force and torque per particle, energy per term. */
struct Acc {
	std::vector<V3> F, Tlab;
	double E[8];
};

/* Each term below follows the reference function of the same name; p and q
 * are already ordered, r = q.pos - p.pos (minimum image). */

double backbone(const oxo_params *P, const Nuc &p, const Nuc &q, V3 r, Acc *A) {
	// src/Interactions/DNAInteraction.cpp:415-466
	V3 rback = r + q.back - p.back;
	double rbackmod = norm(rback);
	double rbackr0 = rbackmod - P->fene_r0;
	V3 force;
	double energy;
	if(P->use_mbf && std::fabs(rbackr0) > P->mbf_xmax) {
		double fene_xmax = -((double) FENE_EPS / 2.f) * std::log(1.f - SQR(P->mbf_xmax) / (double) FENE_DELTA2);
		double long_xmax = (P->mbf_fmax - P->mbf_finf) * P->mbf_xmax * std::log(P->mbf_xmax) + P->mbf_finf * P->mbf_xmax;
		energy = (P->mbf_fmax - P->mbf_finf) * P->mbf_xmax * std::log(std::fabs(rbackr0)) + P->mbf_finf * std::fabs(rbackr0) - long_xmax + fene_xmax;
		force = rback * (-std::copysign(1.f, rbackr0) * ((P->mbf_fmax - P->mbf_finf) * P->mbf_xmax / std::fabs(rbackr0) + P->mbf_finf) / rbackmod);
	}
	else {
		if(std::fabs(rbackr0) > (double) FENE_DELTA - 2.220446049250313e-16) {
			return 1.e12;  // the reference throws when forces are requested
		}
		energy = -((double) FENE_EPS / 2.f) * std::log(1.f - SQR(rbackr0) / (double) FENE_DELTA2);
		force = rback * (-((double) FENE_EPS * rbackr0 / ((double) FENE_DELTA2 - SQR(rbackr0))) / rbackmod);
	}
	if(A) {
		A->F[p.index] -= force;
		A->F[q.index] += force;
		A->Tlab[p.index] -= cross(p.back, force);
		A->Tlab[q.index] += cross(q.back, force);
	}
	return energy;
}

double bonded_excluded_volume(const Nuc &p, const Nuc &q, V3 r, Acc *A) {
	// src/Interactions/DNAInteraction.cpp:468-528
	V3 force, tp{0, 0, 0}, tq{0, 0, 0};
	double energy = repulsive_lj(r + q.base - p.base, force, EXCL_S2, EXCL_R2, EXCL_B2, EXCL_RC2);
	tp -= cross(p.base, force);
	tq += cross(q.base, force);
	V3 Ftot = force;
	energy += repulsive_lj(r + q.back - p.base, force, EXCL_S3, EXCL_R3, EXCL_B3, EXCL_RC3);
	tp -= cross(p.base, force);
	tq += cross(q.back, force);
	Ftot += force;
	energy += repulsive_lj(r + q.base - p.back, force, EXCL_S4, EXCL_R4, EXCL_B4, EXCL_RC4);
	tp -= cross(p.back, force);
	tq += cross(q.base, force);
	Ftot += force;
	if(A) {
		A->F[p.index] -= Ftot;
		A->F[q.index] += Ftot;
		A->Tlab[p.index] += tp;
		A->Tlab[q.index] += tq;
	}
	return energy;
}

double stacking(const oxo_params *P, const Nuc &p, const Nuc &q, V3 r, Acc *A) {
	// src/Interactions/DNAInteraction.cpp:530-705
	const V3 &a1 = p.a1, &a2 = p.a2, &a3 = p.a3, &b1 = q.a1, &b2 = q.a2, &b3 = q.a3;
	V3 rbackref = r + b1 * (double) POS_BACK - a1 * (double) POS_BACK;
	double rbackrefmod = norm(rbackref);
	V3 rstack = r + q.stack - p.stack;
	double rstackmod = norm(rstack);
	V3 rstackdir = rstack / rstackmod;

	double cost4 = dot(a3, b3);
	double cost5 = dot(a3, rstackdir);
	double cost6 = -dot(b3, rstackdir);
	double cosphi1 = dot(a2, rbackref) / rbackrefmod;
	double cosphi2 = dot(b2, rbackref) / rbackrefmod;

	double vf1 = f1(P, rstackmod, F1_STCK, q.type, p.type);
	double f4t4 = custom_f4(cost4, F4_STCK_T4);
	double f4t5 = custom_f4(-cost5, F4_STCK_T5);
	double f4t6 = custom_f4(cost6, F4_STCK_T5);
	double f5phi1 = f5(cosphi1, 0);
	double f5phi2 = f5(cosphi2, 1);
	double energy = vf1 * f4t4 * f4t5 * f4t6 * f5phi1 * f5phi2;

	if(A && energy != 0.) {
		V3 tp{0, 0, 0}, tq{0, 0, 0};
		double vf1D = f1D(P, rstackmod, F1_STCK, q.type, p.type);
		double f4t4Dsin = -custom_f4D(cost4, F4_STCK_T4);
		double f4t5Dsin = -custom_f4D(-cost5, F4_STCK_T5);
		double f4t6Dsin = -custom_f4D(cost6, F4_STCK_T5);
		double f5phi1D = f5D(cosphi1, 0);
		double f5phi2D = f5D(cosphi2, 1);

		V3 force = -rstackdir * (vf1D * f4t4 * f4t5 * f4t6 * f5phi1 * f5phi2);
		force += -(a3 - rstackdir * cost5) * (vf1 * f4t4 * f4t5Dsin * f4t6 * f5phi1 * f5phi2 / rstackmod);
		force += -(b3 + rstackdir * cost6) * (vf1 * f4t4 * f4t5 * f4t6Dsin * f5phi1 * f5phi2 / rstackmod);

		double gamma = (double) POS_STACK - (double) POS_BACK;
		double rbackrefmodcub = rbackrefmod * rbackrefmod * rbackrefmod;

		double ra2 = dot(rstackdir, a2);
		double ra1 = dot(rstackdir, a1);
		double rb1 = dot(rstackdir, b1);
		double a2b1 = dot(a2, b1);
		double parentesi = rstackmod * ra2 - a2b1 * gamma;
		double dcosphi1dr = (SQR(rstackmod) * ra2 - ra2 * SQR(rbackrefmod) - rstackmod * (a2b1 + ra2 * (-ra1 + rb1)) * gamma + a2b1 * (-ra1 + rb1) * SQR(gamma)) / rbackrefmodcub;
		double dcosphi1dra1 = rstackmod * gamma * parentesi / rbackrefmodcub;
		double dcosphi1dra2 = -rstackmod / rbackrefmod;
		double dcosphi1drb1 = -rstackmod * gamma * parentesi / rbackrefmodcub;
		double dcosphi1da1b1 = -SQR(gamma) * parentesi / rbackrefmodcub;
		double dcosphi1da2b1 = gamma / rbackrefmod;
		double force_part_phi1 = -vf1 * f4t4 * f4t5 * f4t6 * f5phi1D * f5phi2;
		force += -(rstackdir * dcosphi1dr + ((a2 - rstackdir * ra2) * dcosphi1dra2 + (a1 - rstackdir * ra1) * dcosphi1dra1 + (b1 - rstackdir * rb1) * dcosphi1drb1) / rstackmod) * force_part_phi1;

		ra2 = dot(rstackdir, b2);
		ra1 = dot(rstackdir, b1);
		rb1 = dot(rstackdir, a1);
		a2b1 = dot(b2, a1);
		parentesi = rstackmod * ra2 + a2b1 * gamma;
		double dcosphi2dr = (parentesi * (rstackmod + (rb1 - ra1) * gamma) - ra2 * SQR(rbackrefmod)) / rbackrefmodcub;
		double dcosphi2dra1 = -rstackmod * gamma * (rstackmod * ra2 + a2b1 * gamma) / rbackrefmodcub;
		double dcosphi2dra2 = -rstackmod / rbackrefmod;
		double dcosphi2drb1 = rstackmod * gamma * parentesi / rbackrefmodcub;
		double dcosphi2da1b1 = -SQR(gamma) * parentesi / rbackrefmodcub;
		double dcosphi2da2b1 = -gamma / rbackrefmod;
		double force_part_phi2 = -vf1 * f4t4 * f4t5 * f4t6 * f5phi1 * f5phi2D;
		force += -force_part_phi2 * (rstackdir * dcosphi2dr + ((b2 - rstackdir * ra2) * dcosphi2dra2 + (b1 - rstackdir * ra1) * dcosphi2dra1 + (a1 - rstackdir * rb1) * dcosphi2drb1) / rstackmod);

		A->F[p.index] -= force;
		A->F[q.index] += force;

		tp -= cross(p.stack, force);
		tq += cross(q.stack, force);

		V3 t4dir = cross(b3, a3);
		double torquemod = vf1 * f4t4Dsin * f4t5 * f4t6 * f5phi1 * f5phi2;
		tp -= t4dir * torquemod;
		tq += t4dir * torquemod;

		V3 t5dir = cross(rstackdir, a3);
		torquemod = -vf1 * f4t4 * f4t5Dsin * f4t6 * f5phi1 * f5phi2;
		tp -= t5dir * torquemod;

		V3 t6dir = cross(rstackdir, b3);
		torquemod = vf1 * f4t4 * f4t5 * f4t6Dsin * f5phi1 * f5phi2;
		tq += t6dir * torquemod;

		tp += cross(rstackdir, a2) * force_part_phi1 * dcosphi1dra2 + cross(rstackdir, a1) * force_part_phi1 * dcosphi1dra1;
		tq += cross(rstackdir, b1) * force_part_phi1 * dcosphi1drb1;
		V3 puretorque = cross(a2, b1) * force_part_phi1 * dcosphi1da2b1 + cross(a1, b1) * force_part_phi1 * dcosphi1da1b1;
		tp -= puretorque;
		tq += puretorque;

		tp += cross(rstackdir, a1) * force_part_phi2 * dcosphi2drb1;
		tq += cross(rstackdir, b2) * force_part_phi2 * dcosphi2dra2 + cross(rstackdir, b1) * force_part_phi2 * dcosphi2dra1;
		puretorque = cross(a1, b2) * force_part_phi2 * dcosphi2da2b1 + cross(a1, b1) * force_part_phi2 * dcosphi2da1b1;
		tp -= puretorque;
		tq += puretorque;

		A->Tlab[p.index] += tp;
		A->Tlab[q.index] += tq;
	}
	return energy;
}

double nonbonded_excluded_volume(const Nuc &p, const Nuc &q, V3 r, Acc *A) {
	// src/Interactions/DNAInteraction.cpp:707-777
	V3 force, tp{0, 0, 0}, tq{0, 0, 0}, Ftot{0, 0, 0};
	double energy = repulsive_lj(r + q.base - p.base, force, EXCL_S2, EXCL_R2, EXCL_B2, EXCL_RC2);
	tp -= cross(p.base, force);
	tq += cross(q.base, force);
	Ftot += force;
	energy += repulsive_lj(r + q.base - p.back, force, EXCL_S4, EXCL_R4, EXCL_B4, EXCL_RC4);
	tp -= cross(p.back, force);
	tq += cross(q.base, force);
	Ftot += force;
	energy += repulsive_lj(r + q.back - p.base, force, EXCL_S3, EXCL_R3, EXCL_B3, EXCL_RC3);
	tp -= cross(p.base, force);
	tq += cross(q.back, force);
	Ftot += force;
	energy += repulsive_lj(r + q.back - p.back, force, EXCL_S1, EXCL_R1, EXCL_B1, EXCL_RC1);
	tp -= cross(p.back, force);
	tq += cross(q.back, force);
	Ftot += force;
	if(A) {
		A->F[p.index] -= Ftot;
		A->F[q.index] += Ftot;
		A->Tlab[p.index] += tp;
		A->Tlab[q.index] += tq;
	}
	return energy;
}

double hydrogen_bonding(const oxo_params *P, const Nuc &p, const Nuc &q, V3 r, Acc *A) {
	// src/Interactions/DNAInteraction.cpp:779-902
	bool is_pair = (q.btype + p.btype == 3);
	double hb_multi = (std::abs(q.btype) >= 300 && std::abs(p.btype) >= 300) ? P->hb_multiplier : 1.;
	V3 rhydro = r + q.base - p.base;
	double rhydromod = norm(rhydro);
	double energy = 0.;
	if(is_pair && (double) HYDR_RCLOW < rhydromod && rhydromod < (double) HYDR_RCHIGH) {
		V3 rhydrodir = rhydro / rhydromod;
		const V3 &a1 = p.a1, &a3 = p.a3, &b1 = q.a1, &b3 = q.a3;
		double cost1 = -dot(a1, b1);
		double cost2 = -dot(b1, rhydrodir);
		double cost3 = dot(a1, rhydrodir);
		double cost4 = dot(a3, b3);
		double cost7 = -dot(b3, rhydrodir);
		double cost8 = dot(a3, rhydrodir);

		double vf1 = hb_multi * f1(P, rhydromod, F1_HYDR, q.type, p.type);
		double f4t1 = custom_f4(cost1, F4_HYDR_T1);
		double f4t2 = custom_f4(cost2, F4_HYDR_T2);
		double f4t3 = custom_f4(cost3, F4_HYDR_T2);
		double f4t4 = custom_f4(cost4, F4_HYDR_T4);
		double f4t7 = custom_f4(cost7, F4_HYDR_T7);
		double f4t8 = custom_f4(cost8, F4_HYDR_T7);
		energy = vf1 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8;

		if(A && energy != 0.) {
			V3 tp{0, 0, 0}, tq{0, 0, 0};
			double vf1D = hb_multi * f1D(P, rhydromod, F1_HYDR, q.type, p.type);
			double f4t1Dsin = custom_f4D(cost1, F4_HYDR_T1);
			double f4t2Dsin = custom_f4D(cost2, F4_HYDR_T2);
			double f4t3Dsin = -custom_f4D(cost3, F4_HYDR_T2);
			double f4t4Dsin = -custom_f4D(cost4, F4_HYDR_T4);
			double f4t7Dsin = custom_f4D(cost7, F4_HYDR_T7);
			double f4t8Dsin = -custom_f4D(cost8, F4_HYDR_T7);

			V3 force = -rhydrodir * (vf1D * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8);

			V3 dir = cross(a3, b3);
			double torquemod = -vf1 * f4t1 * f4t2 * f4t3 * f4t4Dsin * f4t7 * f4t8;
			tp -= dir * torquemod;
			tq += dir * torquemod;

			dir = cross(a1, b1);
			torquemod = -vf1 * f4t1Dsin * f4t2 * f4t3 * f4t4 * f4t7 * f4t8;
			tp -= dir * torquemod;
			tq += dir * torquemod;

			double fact = vf1 * f4t1 * f4t2Dsin * f4t3 * f4t4 * f4t7 * f4t8;
			force += (b1 + rhydrodir * cost2) * (fact / rhydromod);
			dir = cross(rhydrodir, b1);
			tq += -dir * fact;

			force += (a1 - rhydrodir * cost3) * (vf1 * f4t1 * f4t2 * f4t3Dsin * f4t4 * f4t7 * f4t8 / rhydromod);
			V3 t3dir = cross(rhydrodir, a1);
			torquemod = -vf1 * f4t1 * f4t2 * f4t3Dsin * f4t4 * f4t7 * f4t8;
			tp += t3dir * torquemod;

			force += (b3 + rhydrodir * cost7) * (vf1 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7Dsin * f4t8 / rhydromod);
			V3 t7dir = cross(rhydrodir, b3);
			torquemod = -vf1 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7Dsin * f4t8;
			tq += t7dir * torquemod;

			force += (a3 - rhydrodir * cost8) * (vf1 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8Dsin / rhydromod);
			V3 t8dir = cross(rhydrodir, a3);
			torquemod = -vf1 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8Dsin;
			tp += t8dir * torquemod;

			A->F[p.index] -= force;
			A->F[q.index] += force;
			tp -= cross(p.base, force);
			tq += cross(q.base, force);
			A->Tlab[p.index] += tp;
			A->Tlab[q.index] += tq;
		}
	}
	return energy;
}

double cross_stacking(const oxo_params *P, const Nuc &p, const Nuc &q, V3 r, Acc *A) {
	// src/Interactions/DNAInteraction.cpp:904-1021
	V3 rcstack = r + q.base - p.base;
	double rcstackmod = norm(rcstack);
	double energy = 0.;
	if((double) CRST_RCLOW < rcstackmod && rcstackmod < (double) CRST_RCHIGH) {
		V3 rcstackdir = rcstack / rcstackmod;
		const V3 &a1 = p.a1, &a3 = p.a3, &b1 = q.a1, &b3 = q.a3;
		double cost1 = -dot(a1, b1);
		double cost2 = -dot(b1, rcstackdir);
		double cost3 = dot(a1, rcstackdir);
		double cost4 = dot(a3, b3);
		double cost7 = -dot(b3, rcstackdir);
		double cost8 = dot(a3, rcstackdir);

		double vf2 = f2(P, rcstackmod, F2_CRST);
		double f4t1 = custom_f4(cost1, F4_CRST_T1);
		double f4t2 = custom_f4(cost2, F4_CRST_T2);
		double f4t3 = custom_f4(cost3, F4_CRST_T2);
		double f4t4 = custom_f4(cost4, F4_CRST_T4) + custom_f4(-cost4, F4_CRST_T4);
		double f4t7 = custom_f4(cost7, F4_CRST_T7) + custom_f4(-cost7, F4_CRST_T7);
		double f4t8 = custom_f4(cost8, F4_CRST_T7) + custom_f4(-cost8, F4_CRST_T7);
		energy = vf2 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8;

		if(A && energy != 0.) {
			V3 tp{0, 0, 0}, tq{0, 0, 0};
			double vf2D = f2D(P, rcstackmod, F2_CRST);
			double f4t1Dsin = custom_f4D(cost1, F4_CRST_T1);
			double f4t2Dsin = custom_f4D(cost2, F4_CRST_T2);
			double f4t3Dsin = -custom_f4D(cost3, F4_CRST_T2);
			double f4t4Dsin = -custom_f4D(cost4, F4_CRST_T4) + custom_f4D(-cost4, F4_CRST_T4);
			double f4t7Dsin = custom_f4D(cost7, F4_CRST_T7) - custom_f4D(-cost7, F4_CRST_T7);
			double f4t8Dsin = -custom_f4D(cost8, F4_CRST_T7) + custom_f4D(-cost8, F4_CRST_T7);

			V3 force = -rcstackdir * (vf2D * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8);

			V3 t1dir = cross(a1, b1);
			double torquemod = -vf2 * f4t1Dsin * f4t2 * f4t3 * f4t4 * f4t7 * f4t8;
			tp -= t1dir * torquemod;
			tq += t1dir * torquemod;

			force += (b1 + rcstackdir * cost2) * (vf2 * f4t1 * f4t2Dsin * f4t3 * f4t4 * f4t7 * f4t8 / rcstackmod);
			V3 t2dir = cross(rcstackdir, b1);
			torquemod = -vf2 * f4t1 * f4t2Dsin * f4t3 * f4t4 * f4t7 * f4t8;
			tq += t2dir * torquemod;

			force += (a1 - rcstackdir * cost3) * (vf2 * f4t1 * f4t2 * f4t3Dsin * f4t4 * f4t7 * f4t8 / rcstackmod);
			V3 t3dir = cross(rcstackdir, a1);
			torquemod = -vf2 * f4t1 * f4t2 * f4t3Dsin * f4t4 * f4t7 * f4t8;
			tp += t3dir * torquemod;

			V3 t4dir = cross(a3, b3);
			torquemod = -vf2 * f4t1 * f4t2 * f4t3 * f4t4Dsin * f4t7 * f4t8;
			tp -= t4dir * torquemod;
			tq += t4dir * torquemod;

			force += (b3 + rcstackdir * cost7) * (vf2 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7Dsin * f4t8 / rcstackmod);
			V3 t7dir = cross(rcstackdir, b3);
			torquemod = -vf2 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7Dsin * f4t8;
			tq += t7dir * torquemod;

			force += (a3 - rcstackdir * cost8) * (vf2 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8Dsin / rcstackmod);
			V3 t8dir = cross(rcstackdir, a3);
			torquemod = -vf2 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8Dsin;
			tp += t8dir * torquemod;

			A->F[p.index] -= force;
			A->F[q.index] += force;
			tp -= cross(p.base, force);
			tq += cross(q.base, force);
			A->Tlab[p.index] += tp;
			A->Tlab[q.index] += tq;
		}
	}
	return energy;
}

double coaxial_stacking(const oxo_params *P, const Nuc &p, const Nuc &q, V3 r, Acc *A) {
	// oxDNA2 form, src/Interactions/DNA2Interaction.cpp:214-306
	V3 rstack = r + q.stack - p.stack;
	double rstackmod = norm(rstack);
	double energy = 0.;
	if((double) CXST_RCLOW < rstackmod && rstackmod < (double) CXST_RCHIGH) {
		V3 rstackdir = rstack / rstackmod;
		const V3 &a1 = p.a1, &a3 = p.a3, &b1 = q.a1, &b3 = q.a3;
		double cost1 = -dot(a1, b1);
		double cost4 = dot(a3, b3);
		double cost5 = dot(a3, rstackdir);
		double cost6 = -dot(b3, rstackdir);

		double vf2 = f2(P, rstackmod, F2_CXST);
		double f4t1 = custom_f4(cost1, F4_CXST_T1);
		double f4t4 = custom_f4(cost4, F4_CXST_T4);
		double f4t5 = custom_f4(cost5, F4_CXST_T5) + custom_f4(-cost5, F4_CXST_T5);
		double f4t6 = custom_f4(cost6, F4_CXST_T5) + custom_f4(-cost6, F4_CXST_T5);
		energy = vf2 * f4t1 * f4t4 * f4t5 * f4t6;

		if(A && energy != 0.) {
			V3 tp{0, 0, 0}, tq{0, 0, 0};
			double vf2D = f2D(P, rstackmod, F2_CXST);
			double f4t1Dsin = custom_f4D(cost1, F4_CXST_T1);
			double f4t4Dsin = -custom_f4D(cost4, F4_CXST_T4);
			double f4t5Dsin = -custom_f4D(cost5, F4_CXST_T5) + custom_f4D(-cost5, F4_CXST_T5);
			double f4t6Dsin = custom_f4D(cost6, F4_CXST_T5) - custom_f4D(-cost6, F4_CXST_T5);

			V3 force = -rstackdir * (vf2D * f4t1 * f4t4 * f4t5 * f4t6);

			V3 dir = cross(a1, b1);
			double torquemod = -vf2 * f4t1Dsin * f4t4 * f4t5 * f4t6;
			tp -= dir * torquemod;
			tq += dir * torquemod;

			dir = cross(a3, b3);
			torquemod = -vf2 * f4t1 * f4t4Dsin * f4t5 * f4t6;
			tp -= dir * torquemod;
			tq += dir * torquemod;

			double fact = vf2 * f4t1 * f4t4 * f4t5Dsin * f4t6;
			force += fact * (a3 - rstackdir * cost5) / rstackmod;
			dir = cross(rstackdir, a3);
			tp -= dir * fact;

			fact = vf2 * f4t1 * f4t4 * f4t5 * f4t6Dsin;
			force += (b3 + rstackdir * cost6) * (fact / rstackmod);
			dir = cross(rstackdir, b3);
			tq += -dir * fact;

			A->F[p.index] -= force;
			A->F[q.index] += force;
			tp -= cross(p.stack, force);
			tq += cross(q.stack, force);
			A->Tlab[p.index] += tp;
			A->Tlab[q.index] += tq;
		}
	}
	return energy;
}

double debye_huckel(const oxo_params *P, const Nuc &p, const Nuc &q, V3 r, Acc *A) {
	// src/Interactions/DNA2Interaction.cpp:157-212
	double cut_factor = 1.0f;
	if(P->dh_half_charged_ends && (p.n3 < 0 || p.n5 < 0)) cut_factor *= 0.5f;
	if(P->dh_half_charged_ends && (q.n3 < 0 || q.n5 < 0)) cut_factor *= 0.5f;
	V3 rback = r + q.back - p.back;
	double rbackmod = norm(rback);
	double energy = 0.;
	if(rbackmod < P->dh_RC) {
		if(rbackmod < P->dh_RHIGH) energy = std::exp(rbackmod * P->dh_minus_kappa) * (P->dh_prefactor / rbackmod);
		else energy = P->dh_B * SQR(rbackmod - P->dh_RC);
		energy *= cut_factor;
		if(A) {
			V3 force;
			V3 rbackdir = rback / rbackmod;
			if(rbackmod < P->dh_RHIGH) {
				force = rbackdir * (-1.0f * (P->dh_prefactor * std::exp(P->dh_minus_kappa * rbackmod)) * (P->dh_minus_kappa / rbackmod - 1.0f / SQR(rbackmod)));
			}
			else {
				force = -rbackdir * (2.0f * P->dh_B * (rbackmod - P->dh_RC));
			}
			force = force * cut_factor;
			A->F[p.index] -= force;
			A->F[q.index] += force;
			A->Tlab[p.index] += -cross(p.back, force);
			A->Tlab[q.index] += cross(q.back, force);
		}
	}
	return energy;
}

V3 min_image(V3 a, V3 b, const double *box) {
	// src/Boxes/OrthogonalBox.cpp:47-49
	return {b.x - a.x - std::rint((b.x - a.x) / box[0]) * box[0], b.y - a.y - std::rint((b.y - a.y) / box[1]) * box[1],
			b.z - a.z - std::rint((b.z - a.z) / box[2]) * box[2]};
}

std::vector<Nuc> make_nucs(int N, const double *pos, const double *a1, const double *a2, const double *a3, const int *n3, const int *n5,
		const int *btype) {
	std::vector<Nuc> nucs(N);
	for(int i = 0; i < N; i++) {
		Nuc &p = nucs[i];
		p.pos = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
		p.a1 = {a1[3 * i], a1[3 * i + 1], a1[3 * i + 2]};
		p.a2 = {a2[3 * i], a2[3 * i + 1], a2[3 * i + 2]};
		p.a3 = {a3[3 * i], a3[3 * i + 1], a3[3 * i + 2]};
		p.n3 = n3[i];
		p.n5 = n5[i];
		p.btype = btype[i];
		// src/Interactions/DNAInteraction.cpp:1527
		p.type = (p.btype < 0) ? 3 - ((3 - p.btype) % 4) : p.btype % 4;
		p.index = i;
		set_sites(p);
	}
	return nucs;
}

/* Unordered candidate pairs (i > j) within rcut, by a cell list for large N
 * and brute force otherwise.  The candidate set is a superset of the pairs
 * with a non-zero interaction; the terms themselves are range limited. */
void candidate_pairs(const std::vector<Nuc> &nucs, const double *box, double rc, std::vector<std::pair<int, int>> &out) {
	int N = (int) nucs.size();
	double rc2 = rc * rc;
	out.clear();
	int nc[3];
	for(int d = 0; d < 3; d++) nc[d] = (int) std::floor(box[d] / rc);
	if(N < 1500 || nc[0] < 3 || nc[1] < 3 || nc[2] < 3) {
		for(int i = 0; i < N; i++)
			for(int j = 0; j < i; j++) {
				V3 r = min_image(nucs[i].pos, nucs[j].pos, box);
				if(dot(r, r) < rc2) out.push_back({i, j});
			}
		return;
	}
	int ncell = nc[0] * nc[1] * nc[2];
	std::vector<int> head(ncell, -1), next(N, -1), cidx(N);
	for(int i = N - 1; i >= 0; i--) {
		int c[3];
		double p[3] = {nucs[i].pos.x, nucs[i].pos.y, nucs[i].pos.z};
		for(int d = 0; d < 3; d++) {
			double f = p[d] / box[d] - std::floor(p[d] / box[d]);
			c[d] = std::min(nc[d] - 1, (int) (f * nc[d]));
		}
		cidx[i] = (c[2] * nc[1] + c[1]) * nc[0] + c[0];
		next[i] = head[cidx[i]];
		head[cidx[i]] = i;
	}
	for(int i = 0; i < N; i++) {
		int c0 = cidx[i] % nc[0], c1 = (cidx[i] / nc[0]) % nc[1], c2 = cidx[i] / (nc[0] * nc[1]);
		for(int dz = -1; dz <= 1; dz++)
			for(int dy = -1; dy <= 1; dy++)
				for(int dx = -1; dx <= 1; dx++) {
					int cc = (((c2 + dz + nc[2]) % nc[2]) * nc[1] + (c1 + dy + nc[1]) % nc[1]) * nc[0] + (c0 + dx + nc[0]) % nc[0];
					for(int j = head[cc]; j >= 0; j = next[j]) {
						if(j >= i) continue;
						V3 r = min_image(nucs[i].pos, nucs[j].pos, box);
						if(dot(r, r) < rc2) out.push_back({i, j});
					}
				}
	}
}

bool is_bonded(const Nuc &p, const Nuc &q) { return p.n3 == q.index || p.n5 == q.index; }

}  // namespace

extern "C" {

/* Model constants for one system; mirrors DNAInteraction::init
 * (src/Interactions/DNAInteraction.cpp:298-326) and DNA2Interaction's
 * get_settings / init (src/Interactions/DNA2Interaction.cpp:62-155).
 * Negative dh_lambda / dh_strength / dh_rhigh / max_backbone_force mean
 * "key absent".  The salt concentration is held as float, as the
 * reference's DNA2Interaction::_salt_concentration is. */
int oxo_params_init(oxo_params *P, double T, double salt, int half_charged_ends, double dh_lambda, double dh_strength, double dh_rhigh,
		double hb_multiplier, double max_backbone_force, double max_backbone_force_far) {
	std::memset(P, 0, sizeof(*P));
	P->T = T;
	const float salt_f = (float) salt;
	P->salt = salt_f;
	P->hb_multiplier = hb_multiplier;
	P->dh_half_charged_ends = half_charged_ends;
	P->fene_r0 = (double) FENE_R0_OXDNA2;

	P->F1_A[0] = HYDR_A; P->F1_A[1] = STCK_A;
	P->F1_RC[0] = HYDR_RC; P->F1_RC[1] = STCK_RC;
	P->F1_R0[0] = HYDR_R0; P->F1_R0[1] = STCK_R0;
	P->F1_BLOW[0] = HYDR_BLOW; P->F1_BLOW[1] = STCK_BLOW;
	P->F1_BHIGH[0] = HYDR_BHIGH; P->F1_BHIGH[1] = STCK_BHIGH;
	P->F1_RLOW[0] = HYDR_RLOW; P->F1_RLOW[1] = STCK_RLOW;
	P->F1_RHIGH[0] = HYDR_RHIGH; P->F1_RHIGH[1] = STCK_RHIGH;
	P->F1_RCLOW[0] = HYDR_RCLOW; P->F1_RCLOW[1] = STCK_RCLOW;
	P->F1_RCHIGH[0] = HYDR_RCHIGH; P->F1_RCHIGH[1] = STCK_RCHIGH;

	P->F2_K[0] = CRST_K; P->F2_K[1] = CXST_K_OXDNA2;
	P->F2_RC[0] = CRST_RC; P->F2_RC[1] = CXST_RC;
	P->F2_R0[0] = CRST_R0; P->F2_R0[1] = CXST_R0;
	P->F2_BLOW[0] = CRST_BLOW; P->F2_BLOW[1] = CXST_BLOW;
	P->F2_BHIGH[0] = CRST_BHIGH; P->F2_BHIGH[1] = CXST_BHIGH;
	P->F2_RLOW[0] = CRST_RLOW; P->F2_RLOW[1] = CXST_RLOW;
	P->F2_RHIGH[0] = CRST_RHIGH; P->F2_RHIGH[1] = CXST_RHIGH;
	P->F2_RCLOW[0] = CRST_RCLOW; P->F2_RCLOW[1] = CXST_RCLOW;
	P->F2_RCHIGH[0] = CRST_RCHIGH; P->F2_RCHIGH[1] = CXST_RCHIGH;

	// average-sequence well depths (src/Interactions/DNA2Interaction.cpp:112-124)
	for(int i = 0; i < 5; i++)
		for(int j = 0; j < 5; j++) {
			P->F1_EPS[F1_STCK][i][j] = (double) STCK_BASE_EPS_OXDNA2 + (double) STCK_FACT_EPS_OXDNA2 * T;
			P->F1_SHIFT[F1_STCK][i][j] = P->F1_EPS[F1_STCK][i][j] * SQR(1 - std::exp((double) (-(STCK_RC - STCK_R0) * STCK_A)));
			P->F1_EPS[F1_HYDR][i][j] = (double) HYDR_EPS_OXDNA2;
			P->F1_SHIFT[F1_HYDR][i][j] = P->F1_EPS[F1_HYDR][i][j] * SQR(1 - std::exp((double) (-(HYDR_RC - HYDR_R0) * HYDR_A)));
		}

	// Debye-Hueckel (src/Interactions/DNA2Interaction.cpp:68-81,126-147)
	P->dh_lambdafactor = dh_lambda >= 0 ? dh_lambda : DH_LAMBDA_DEFAULT;
	P->dh_prefactor = dh_strength >= 0 ? dh_strength : DH_STRENGTH_DEFAULT;
	// the float exponent arguments and the float salt go through the FP64
	// library calls, as in the reference
	double lambda_settings = P->dh_lambdafactor * std::sqrt(T / 0.1f) / std::sqrt((double) salt_f);
	P->dh_RHIGH = dh_rhigh > 0 ? dh_rhigh : 3.0 * lambda_settings;
	double lambda = P->dh_lambdafactor * std::sqrt(T / 0.1) / std::sqrt((double) salt_f);
	P->dh_minus_kappa = -1.0 / lambda;
	double x = P->dh_RHIGH, q = P->dh_prefactor, l = lambda;
	P->dh_B = -(std::exp(-x / l) * q * q * (x + l) * (x + l)) / (-4. * x * x * x * l * l * q);
	P->dh_RC = x * (q * x + 3. * q * l) / (q * (x + l));

	// cut-off (src/Interactions/DNAInteraction.cpp:301-313, DNA2Interaction.cpp:139-152)
	double back_arm = std::sqrt(SQR((double) POS_MM_BACK1) + SQR((double) POS_MM_BACK2));
	double rcutback = 2 * back_arm + (double) EXCL_RC1;
	double rcutbase = 2 * std::fabs((double) POS_BASE) + (double) HYDR_RCHIGH;
	P->rcut = std::fmax(rcutback, rcutbase);
	double debyecut = 2.0 * back_arm + P->dh_RC;
	if(debyecut > P->rcut) P->rcut = debyecut;
	P->sqr_rcut = SQR(P->rcut);

	// max_backbone_force (src/Interactions/DNAInteraction.cpp:271-287)
	if(max_backbone_force >= 0) {
		P->use_mbf = 1;
		P->mbf_fmax = max_backbone_force;
		P->mbf_xmax = (-(double) FENE_EPS + std::sqrt((double) FENE_EPS * (double) FENE_EPS + 4.f * P->mbf_fmax * P->mbf_fmax * (double) FENE_DELTA2)) / (2.f * P->mbf_fmax);
		P->mbf_finf = max_backbone_force_far >= 0 ? max_backbone_force_far : 0.04f;
	}
	return 0;
}

/* Overrides the stacking and H-bond well depths with sequence-dependent
 * values, as DNAInteraction::init does from seq_dep_file
 * (src/Interactions/DNAInteraction.cpp:328-376): stck[i][j] for bases
 * i, j in {A=0, G=1, C=2, T=3}, hydr_AT, hydr_GC. */
int oxo_params_set_seqdep(oxo_params *P, const double *stck16, double stck_fact_eps, double hydr_AT, double hydr_GC) {
	for(int i = 0; i < 4; i++)
		for(int j = 0; j < 4; j++) {
			P->F1_EPS[F1_STCK][i][j] = stck16[4 * i + j] * (1.0 - stck_fact_eps + (P->T * 9.0 * stck_fact_eps));
			P->F1_SHIFT[F1_STCK][i][j] = P->F1_EPS[F1_STCK][i][j] * SQR(1 - std::exp((double) (-(STCK_RC - STCK_R0) * STCK_A)));
		}
	const int A = 0, G = 1, C = 2, T = 3;
	P->F1_EPS[F1_HYDR][A][T] = P->F1_EPS[F1_HYDR][T][A] = hydr_AT;
	P->F1_EPS[F1_HYDR][G][C] = P->F1_EPS[F1_HYDR][C][G] = hydr_GC;
	double sh = SQR(1 - std::exp((double) (-(HYDR_RC - HYDR_R0) * HYDR_A)));
	P->F1_SHIFT[F1_HYDR][A][T] = P->F1_SHIFT[F1_HYDR][T][A] = hydr_AT * sh;
	P->F1_SHIFT[F1_HYDR][G][C] = P->F1_SHIFT[F1_HYDR][C][G] = hydr_GC * sh;
	return 0;
}

int oxo_params_size(void) { return (int) sizeof(oxo_params); }
double oxo_params_rcut(const oxo_params *P) { return P->rcut; }
void oxo_params_dh(const oxo_params *P, double *out4) {
	out4[0] = P->dh_RHIGH;
	out4[1] = P->dh_RC;
	out4[2] = P->dh_B;
	out4[3] = P->dh_minus_kappa;
}

/* Energies, forces and body-frame torques of a whole system, as the
 * reference's MD_CPUBackend::_compute_forces evaluates them
 * (src/Backends/MD_CPUBackend.cpp:146-167): every bonded pair once with the
 * 5'-side particle first (DNAInteraction::_check_bonded_neighbour,
 * src/Interactions/DNAInteraction.cpp:389-413), every non-bonded pair once
 * with centre distance below rcut.
 *   E8 (out, may be NULL): total energy of each term, in the order of the
 *      reference's energy split (FENE, bonded excl, stacking, non-bonded
 *      excl, HB, CRST, CXST, DH; src/Interactions/DNA2Interaction.h:93-95).
 *   F, Tbody (out, 3N each, may be NULL).
 *   E_part (out, N, may be NULL): per-particle energy, each pair's energy
 *      split half-half, so that sum(E_part) = total energy.
 * Returns the total potential energy. */
static double forces_impl(const oxo_params *P, int N, const double *pos, const double *a1, const double *a2, const double *a3,
		const int *n3, const int *n5, const int *btype, const double *box, double *E8, double *F, double *Tbody, double *E_part,
		double *W6);

double oxo_forces(const oxo_params *P, int N, const double *pos, const double *a1, const double *a2, const double *a3, const int *n3,
		const int *n5, const int *btype, const double *box, double *E8, double *F, double *Tbody, double *E_part) {
	return forces_impl(P, N, pos, a1, a2, a3, n3, n5, btype, box, E8, F, Tbody, E_part, nullptr);
}

/* Pair virial of the stress tensor that the DNA interactions accumulate
 * during a force computation (DNAInteraction's _update_stress_tensor
 * calls, src/Interactions/BaseInteraction.cpp:159-200; the pressure
 * observable divides by the volume and adds the kinetic term,
 * BaseInteraction.cpp:285-352): sum over every interacting pair (bonded and
 * non-bonded, once) of -r (x) F_p, r = q - p the nearest-image centre
 * separation and F_p the pair's total force on p; components xx, yy, zz,
 * xy, xz, yz.  Returns the potential energy. */
double oxo_virial(const oxo_params *P, int N, const double *pos, const double *a1, const double *a2, const double *a3, const int *n3,
		const int *n5, const int *btype, const double *box, double *W6) {
	for(int k = 0; k < 6; k++) W6[k] = 0.;
	return forces_impl(P, N, pos, a1, a2, a3, n3, n5, btype, box, nullptr, nullptr, nullptr, nullptr, W6);
}

static void add_virial(double *W, const V3 &r, const V3 &dFp) {
	if(!W) return;
	W[0] -= r.x * dFp.x;
	W[1] -= r.y * dFp.y;
	W[2] -= r.z * dFp.z;
	W[3] -= r.x * dFp.y;
	W[4] -= r.x * dFp.z;
	W[5] -= r.y * dFp.z;
}

static double forces_impl(const oxo_params *P, int N, const double *pos, const double *a1, const double *a2, const double *a3,
		const int *n3, const int *n5, const int *btype, const double *box, double *E8, double *F, double *Tbody, double *E_part,
		double *W6) {
	std::vector<Nuc> nucs = make_nucs(N, pos, a1, a2, a3, n3, n5, btype);
	Acc A;
	A.F.assign(N, V3{0, 0, 0});
	A.Tlab.assign(N, V3{0, 0, 0});
	for(double &e : A.E) e = 0.;
	std::vector<double> ep(N, 0.);

	// bonded pairs, p -> n3 = q
	for(int i = 0; i < N; i++) {
		if(nucs[i].n3 < 0) continue;
		const Nuc &p = nucs[i];
		const Nuc &q = nucs[p.n3];
		V3 r = min_image(p.pos, q.pos, box);
		const V3 F0 = A.F[p.index];
		double e0 = backbone(P, p, q, r, &A);
		double e1 = bonded_excluded_volume(p, q, r, &A);
		double e2 = stacking(P, p, q, r, &A);
		add_virial(W6, r, A.F[p.index] - F0);
		A.E[0] += e0;
		A.E[1] += e1;
		A.E[2] += e2;
		ep[p.index] += 0.5 * (e0 + e1 + e2);
		ep[q.index] += 0.5 * (e0 + e1 + e2);
	}

	// non-bonded pairs, the higher index first (half list as in src/Lists/Cells.cpp:168)
	std::vector<std::pair<int, int>> pairs;
	candidate_pairs(nucs, box, P->rcut, pairs);
	for(auto &pr : pairs) {
		const Nuc &p = nucs[pr.first];
		const Nuc &q = nucs[pr.second];
		if(is_bonded(p, q)) continue;
		V3 r = min_image(p.pos, q.pos, box);
		if(dot(r, r) >= P->sqr_rcut) continue;
		const V3 F0 = A.F[p.index];
		double e3 = nonbonded_excluded_volume(p, q, r, &A);
		double e4 = hydrogen_bonding(P, p, q, r, &A);
		double e5 = cross_stacking(P, p, q, r, &A);
		double e6 = coaxial_stacking(P, p, q, r, &A);
		double e7 = debye_huckel(P, p, q, r, &A);
		add_virial(W6, r, A.F[p.index] - F0);
		A.E[3] += e3;
		A.E[4] += e4;
		A.E[5] += e5;
		A.E[6] += e6;
		A.E[7] += e7;
		double es = e3 + e4 + e5 + e6 + e7;
		ep[p.index] += 0.5 * es;
		ep[q.index] += 0.5 * es;
	}

	double tot = 0.;
	for(int k = 0; k < 8; k++) {
		if(E8) E8[k] = A.E[k];
		tot += A.E[k];
	}
	for(int i = 0; i < N; i++) {
		if(F) {
			F[3 * i] = A.F[i].x;
			F[3 * i + 1] = A.F[i].y;
			F[3 * i + 2] = A.F[i].z;
		}
		if(Tbody) {
			// orientationT * T_lab
			const Nuc &p = nucs[i];
			Tbody[3 * i] = dot(p.a1, A.Tlab[i]);
			Tbody[3 * i + 1] = dot(p.a2, A.Tlab[i]);
			Tbody[3 * i + 2] = dot(p.a3, A.Tlab[i]);
		}
		if(E_part) E_part[i] = ep[i];
	}
	return tot;
}

/* Per-pair HB energies of all complementary pairs (used to count base
 * pairs with the reference's HB_CUTOFF, src/Observables/HBList.cpp:141-164).
 * Returns the number of pairs with E_HB < HB_CUTOFF. */
int oxo_count_hb(const oxo_params *P, int N, const double *pos, const double *a1, const double *a2, const double *a3, const int *n3,
		const int *n5, const int *btype, const double *box) {
	std::vector<Nuc> nucs = make_nucs(N, pos, a1, a2, a3, n3, n5, btype);
	std::vector<std::pair<int, int>> pairs;
	candidate_pairs(nucs, box, P->rcut, pairs);
	int count = 0;
	for(auto &pr : pairs) {
		const Nuc &p = nucs[pr.first];
		const Nuc &q = nucs[pr.second];
		if(is_bonded(p, q)) continue;
		V3 r = min_image(p.pos, q.pos, box);
		if(hydrogen_bonding(P, p, q, r, nullptr) < (double) HB_CUTOFF) count++;
	}
	return count;
}

/* ------------------------------------------------------------------------
 * Verlet pair predicate.  The pair set of the HIP list is defined by this
 * exact FP64 arithmetic on the FP64 positions: minimum image with rint of a
 * product by the host-computed 1/L, explicit fused multiply-adds.  Because
 * IEEE FP64 fma / mul / sub / rint are correctly rounded on both sides, the
 * GPU and this function agree bit for bit.  The set itself is the
 * reference's: all j != i, not bonded to i, with |r_ij|^2 < r_verlet^2
 * (src/CUDA/Lists/CUDA_simple_verlet.cuh:23-38).
 * Writes pairs (i < j) into out (2 ints each) up to max_pairs; returns the
 * number of pairs (which may exceed max_pairs). */
long long oxo_verlet_pairs(int N, const double *pos, const int *n3, const int *n5, const double *box, double sqr_rverlet, int *out,
		long long max_pairs) {
	double inv[3] = {1.0 / box[0], 1.0 / box[1], 1.0 / box[2]};
	double rv = std::sqrt(sqr_rverlet);
	int nc[3];
	for(int d = 0; d < 3; d++) nc[d] = std::max(1, (int) std::floor(box[d] / rv));
	bool brute = N < 2000 || nc[0] < 3 || nc[1] < 3 || nc[2] < 3;
	auto r2 = [&](int i, int j) {
		double dx = pos[3 * j] - pos[3 * i];
		double dy = pos[3 * j + 1] - pos[3 * i + 1];
		double dz = pos[3 * j + 2] - pos[3 * i + 2];
		dx = std::fma(-box[0], std::rint(dx * inv[0]), dx);
		dy = std::fma(-box[1], std::rint(dy * inv[1]), dy);
		dz = std::fma(-box[2], std::rint(dz * inv[2]), dz);
		return std::fma(dz, dz, std::fma(dy, dy, dx * dx));
	};
	long long count = 0;
	auto emit = [&](int i, int j) {
		if(n3[i] == j || n5[i] == j) return;
		if(r2(i, j) < sqr_rverlet) {
			if(count < max_pairs) {
				out[2 * count] = std::min(i, j);
				out[2 * count + 1] = std::max(i, j);
			}
			count++;
		}
	};
	if(brute) {
		for(int i = 0; i < N; i++)
			for(int j = i + 1; j < N; j++) emit(i, j);
		return count;
	}
	int ncell = nc[0] * nc[1] * nc[2];
	std::vector<int> head(ncell, -1), next(N, -1), cidx(N);
	for(int i = N - 1; i >= 0; i--) {
		int c[3];
		for(int d = 0; d < 3; d++) {
			double s = pos[3 * i + d] * inv[d];
			c[d] = std::min(nc[d] - 1, (int) ((s - std::floor(s)) * nc[d]));
		}
		cidx[i] = (c[2] * nc[1] + c[1]) * nc[0] + c[0];
		next[i] = head[cidx[i]];
		head[cidx[i]] = i;
	}
	for(int i = 0; i < N; i++) {
		int c0 = cidx[i] % nc[0], c1 = (cidx[i] / nc[0]) % nc[1], c2 = cidx[i] / (nc[0] * nc[1]);
		for(int dz = -1; dz <= 1; dz++)
			for(int dy = -1; dy <= 1; dy++)
				for(int dx = -1; dx <= 1; dx++) {
					int cc = (((c2 + dz + nc[2]) % nc[2]) * nc[1] + (c1 + dy + nc[1]) % nc[1]) * nc[0] + (c0 + dx + nc[0]) % nc[0];
					for(int j = head[cc]; j >= 0; j = next[j])
						if(j > i) emit(i, j);
				}
	}
	return count;
}

/* ------------------------------------------------------------------------
 * Hilbert key of a cell (x, y, z), each coordinate on `bits` bits, after
 * J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707, 381
 * (2004): axes -> transposed index (inverse undo + Gray encode), then the
 * transposed bits interleaved most-significant first.  The reference
 * orders particles along a Hilbert curve as well (after Wang & Shan 2005,
 * src/CUDA/CUDA_sort.cu:37-104); any Hilbert curve serves the purpose
 * (locality), and this one is pinned in tests by bijectivity and by
 * consecutive keys mapping to face-adjacent cells. */
unsigned int oxo_hilbert_key(unsigned int x, unsigned int y, unsigned int z, int bits) {
	unsigned int X[3] = {x, y, z};
	unsigned int M = 1u << (bits - 1);
	for(unsigned int Q = M; Q > 1; Q >>= 1) {
		unsigned int P = Q - 1;
		for(int i = 0; i < 3; i++) {
			if(X[i] & Q) X[0] ^= P;
			else {
				unsigned int t = (X[0] ^ X[i]) & P;
				X[0] ^= t;
				X[i] ^= t;
			}
		}
	}
	for(int i = 1; i < 3; i++) X[i] ^= X[i - 1];
	unsigned int t = 0;
	for(unsigned int Q = M; Q > 1; Q >>= 1)
		if(X[2] & Q) t ^= Q - 1;
	for(int i = 0; i < 3; i++) X[i] ^= t;
	unsigned int key = 0;
	for(int b = bits - 1; b >= 0; b--)
		for(int i = 0; i < 3; i++) key = (key << 1) | ((X[i] >> b) & 1u);
	return key;
}

/* ------------------------------------------------------------------------
 * Philox4x32-10 counter-based generator (J. K. Salmon, M. A. Moraes,
 * R. O. Dror, D. E. Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11).  ctr[4], key[2] in; 4 words out. */
void oxo_philox4x32_10(const unsigned int *ctr_in, const unsigned int *key_in, unsigned int *out) {
	uint32_t c[4] = {ctr_in[0], ctr_in[1], ctr_in[2], ctr_in[3]};
	uint32_t k[2] = {key_in[0], key_in[1]};
	for(int r = 0; r < 10; r++) {
		uint64_t p0 = (uint64_t) 0xD2511F53u * c[0];
		uint64_t p1 = (uint64_t) 0xCD9E8D57u * c[2];
		uint32_t hi0 = (uint32_t) (p0 >> 32), lo0 = (uint32_t) p0;
		uint32_t hi1 = (uint32_t) (p1 >> 32), lo1 = (uint32_t) p1;
		uint32_t n0 = hi1 ^ c[1] ^ k[0];
		uint32_t n2 = hi0 ^ c[3] ^ k[1];
		c[0] = n0;
		c[1] = lo1;
		c[2] = n2;
		c[3] = lo0;
		k[0] += 0x9E3779B9u;
		k[1] += 0xBB67AE85u;
	}
	for(int i = 0; i < 4; i++) out[i] = c[i];
}

/* ------------------------------------------------------------------------
 * Integrator: one velocity-Verlet step with unit mass and inertia, as
 * MD_CPUBackend::_first_step / _second_step (src/Backends/MD_CPUBackend.cpp:
 * 67-144,169-180), the orientation rotated by the Rodrigues matrix of the
 * body-frame angular momentum.  State arrays are updated in place; F and
 * Tbody must hold the forces of the current state on entry and hold the new
 * forces on exit.  Returns the potential energy after the step. */
double oxo_md_step(const oxo_params *P, int N, double *pos, double *a1, double *a2, double *a3, double *vel, double *L, const int *n3,
		const int *n5, const int *btype, const double *box, double dt, double *F, double *Tbody) {
	for(int i = 0; i < N; i++) {
		for(int d = 0; d < 3; d++) {
			vel[3 * i + d] += F[3 * i + d] * (dt * 0.5);
			pos[3 * i + d] += vel[3 * i + d] * dt;
		}
		V3 Lv{L[3 * i], L[3 * i + 1], L[3 * i + 2]};
		Lv += V3{Tbody[3 * i], Tbody[3 * i + 1], Tbody[3 * i + 2]} * (dt * 0.5);
		L[3 * i] = Lv.x;
		L[3 * i + 1] = Lv.y;
		L[3 * i + 2] = Lv.z;
		double nrm = norm(Lv);
		V3 u = Lv / nrm;
		double s = std::sin(dt * nrm), c = std::cos(dt * nrm), oc = 1. - c;
		double R[3][3] = {{u.x * u.x * oc + c, u.x * u.y * oc - u.z * s, u.x * u.z * oc + u.y * s},
				{u.x * u.y * oc + u.z * s, u.y * u.y * oc + c, u.y * u.z * oc - u.x * s},
				{u.x * u.z * oc - u.y * s, u.y * u.z * oc + u.x * s, u.z * u.z * oc + c}};
		// orientation O has columns a1, a2, a3; O' = O R
		V3 A[3] = {{a1[3 * i], a1[3 * i + 1], a1[3 * i + 2]}, {a2[3 * i], a2[3 * i + 1], a2[3 * i + 2]}, {a3[3 * i], a3[3 * i + 1], a3[3 * i + 2]}};
		V3 An[3];
		for(int k = 0; k < 3; k++) An[k] = A[0] * R[0][k] + A[1] * R[1][k] + A[2] * R[2][k];
		for(int d = 0; d < 3; d++) {
			a1[3 * i + d] = (&An[0].x)[d];
			a2[3 * i + d] = (&An[1].x)[d];
			a3[3 * i + d] = (&An[2].x)[d];
		}
	}
	double U = oxo_forces(P, N, pos, a1, a2, a3, n3, n5, btype, box, nullptr, F, Tbody, nullptr);
	for(int i = 0; i < N; i++)
		for(int d = 0; d < 3; d++) {
			vel[3 * i + d] += F[3 * i + d] * dt * 0.5;
			L[3 * i + d] += Tbody[3 * i + d] * dt * 0.5;
		}
	return U;
}

}  // extern "C"
