// Pair physics of the oxDNA2 interaction on gfx950, shared by the grid
// force kernels (oxh_forces.hip) and the workgroup-resident engine
// (oxh_persist.hip).
//
// Terms follow the reference's CPU DNA2_nomesh formulation
// (src/Interactions/DNAInteraction.cpp:415-1205,
//  src/Interactions/DNA2Interaction.cpp:157-306), in FP32.
#pragma once

#include "oxh_device.h"

namespace oxh {

struct Frame {
	f3 a1, a2, a3, back, stack, base;
	int btype, type;
};

__device__ __forceinline__ int type_of(int btype) { return (btype < 0) ? 3 - ((3 - btype) % 4) : btype % 4; }

__device__ __forceinline__ Frame make_frame(float4 q, int btype) {
	Frame fr;
	quat_axes(q, fr.a1, fr.a2, fr.a3);
	fr.back = fr.a1 * POS_MM_BACK1 + fr.a2 * POS_MM_BACK2;
	fr.stack = fr.a1 * POS_STACK;
	fr.base = fr.a1 * POS_BASE;
	fr.btype = btype;
	fr.type = type_of(btype);
	return fr;
}

__device__ __forceinline__ f3 min_image(double4 a, double4 b, const oxh_sys &S) {
	double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
	dx = fma(-S.box[0], rint(dx * S.inv_box[0]), dx);
	dy = fma(-S.box[1], rint(dy * S.inv_box[1]), dy);
	dz = fma(-S.box[2], rint(dz * S.inv_box[2]), dz);
	return mk((float) dx, (float) dy, (float) dz);
}

// Result of one pair term evaluated for the ordered pair (p, q), r = q - p:
// force on q (p receives -force), torques on p and q (lab frame), energy.
struct PairOut {
	f3 force, tp, tq;
	float e;
};

// ------------------------------------------------ FP64 site geometry
// The pair terms are FP32, but the site-to-site separations they start
// from are formed in FP64 and rounded once: the radial factors (FENE,
// Morse, harmonic, the steep excluded-volume walls) depend on
// |r_site| - r0, which is small next to |r_site|, so FP32 sites would lose
// ~1e-6 of the force to cancellation.  Axes come from the FP64
// quaternion (as the FP32 mirror forms them) and are rounded once.
struct d3 {
	double x, y, z;
};

// a nucleotide's FP64 a1 axis and backbone offset (the stacking and base
// offsets are POS_STACK a1 and POS_BASE a1)
struct Sites64 {
	d3 a1, back;
};

__device__ __forceinline__ Frame make_frame_d(double4 q, int btype, Sites64 &st) {
	const double sqx = q.x * q.x, sqy = q.y * q.y, sqz = q.z * q.z, sqw = q.w * q.w;
	const double xy = q.x * q.y, xz = q.x * q.z, xw = q.x * q.w, yz = q.y * q.z, yw = q.y * q.w, zw = q.z * q.w;
	const d3 a1 = {sqx - sqy - sqz + sqw, 2. * (xy + zw), 2. * (xz - yw)};
	const d3 a2 = {2. * (xy - zw), -sqx + sqy - sqz + sqw, 2. * (yz + xw)};
	const d3 a3 = {2. * (xz + yw), 2. * (yz - xw), -sqx - sqy + sqz + sqw};
	const double b1 = (double) POS_MM_BACK1, b2 = (double) POS_MM_BACK2, ps = (double) POS_STACK, pb = (double) POS_BASE;
	st.a1 = a1;
	st.back = {b1 * a1.x + b2 * a2.x, b1 * a1.y + b2 * a2.y, b1 * a1.z + b2 * a2.z};
	Frame fr;
	fr.a1 = mk((float) a1.x, (float) a1.y, (float) a1.z);
	fr.a2 = mk((float) a2.x, (float) a2.y, (float) a2.z);
	fr.a3 = mk((float) a3.x, (float) a3.y, (float) a3.z);
	fr.back = mk((float) st.back.x, (float) st.back.y, (float) st.back.z);
	fr.stack = mk((float) (ps * a1.x), (float) (ps * a1.y), (float) (ps * a1.z));
	fr.base = mk((float) (pb * a1.x), (float) (pb * a1.y), (float) (pb * a1.z));
	fr.btype = btype;
	fr.type = type_of(btype);
	return fr;
}

__device__ __forceinline__ d3 min_image_d(double4 a, double4 b, const oxh_sys &S) {
	double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
	dx = fma(-S.box[0], rint(dx * S.inv_box[0]), dx);
	dy = fma(-S.box[1], rint(dy * S.inv_box[1]), dy);
	dz = fma(-S.box[2], rint(dz * S.inv_box[2]), dz);
	return {dx, dy, dz};
}

// (r + qs) - ps in FP64, rounded to FP32
__device__ __forceinline__ f3 site_sep(const d3 &r, const d3 &qs, const d3 &ps) {
	return mk((float) ((r.x + qs.x) - ps.x), (float) ((r.y + qs.y) - ps.y), (float) ((r.z + qs.z) - ps.z));
}

// separations of the site pairs the terms use, for the ordered pair
// (p, q), r = q - p
struct PairGeom {
	f3 r;          // centres
	f3 back;       // back(q) - back(p)
	f3 base;       // base(q) - base(p)
	f3 base_back;  // base(q) - back(p)
	f3 back_base;  // back(q) - base(p)
	f3 stack;      // stack(q) - stack(p)
};

__device__ __forceinline__ d3 scaled(const d3 &v, double c) { return {c * v.x, c * v.y, c * v.z}; }

__device__ __forceinline__ PairGeom pair_geom(const d3 &r, const Sites64 &p, const Sites64 &q) {
	const double pb = (double) POS_BASE, ps = (double) POS_STACK;
	const d3 pbase = scaled(p.a1, pb), qbase = scaled(q.a1, pb);
	PairGeom g;
	g.r = mk((float) r.x, (float) r.y, (float) r.z);
	g.back = site_sep(r, q.back, p.back);
	g.base = site_sep(r, qbase, pbase);
	g.base_back = site_sep(r, qbase, p.back);
	g.back_base = site_sep(r, q.back, pbase);
	g.stack = site_sep(r, scaled(q.a1, ps), scaled(p.a1, ps));
	return g;
}

// bonded pairs also need the FENE stretch |back(q) - back(p)| - r0 (FP64)
// and the separation of the POS_BACK reference sites of the stacking's
// phi angles
__device__ __forceinline__ void bonded_geom(const d3 &r, const Sites64 &p, const Sites64 &q, const oxh_sys &S, float &stretch,
		f3 &backref) {
	const double dx = (r.x + q.back.x) - p.back.x, dy = (r.y + q.back.y) - p.back.y, dz = (r.z + q.back.z) - p.back.z;
	stretch = (float) (sqrt(dx * dx + dy * dy + dz * dz) - (double) S.fene_r0);
	const double pbk = (double) POS_BACK;
	backref = mk((float) (r.x + pbk * (q.a1.x - p.a1.x)), (float) (r.y + pbk * (q.a1.y - p.a1.y)),
			(float) (r.z + pbk * (q.a1.z - p.a1.z)));
}

// ------------------------------------------------------------ bonded terms
// FENE, bonded excluded volume and stacking of the bonded pair p -> n3 = q
// (src/Interactions/DNAInteraction.cpp:415-705); stretch = fene_stretch of
// the pair; returns the three energies in e3.
__device__ PairOut bonded_pair(const oxh_sys &S, const Frame &p, const Frame &q, const PairGeom &g, float stretch, f3 rbackref,
		float e3[3], int &err) {
	PairOut o = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, 0.f};

	// FENE backbone
	f3 rback = g.back;
	float rbackr0 = stretch;
	float rbackmod = S.fene_r0 + stretch;
	f3 force;
	float e_fene;
	if(S.use_mbf && fabsf(rbackr0) > S.mbf_xmax) {
		float fene_xmax = -(FENE_EPS * 0.5f) * logf(1.f - SQR(S.mbf_xmax) / FENE_DELTA2);
		float long_xmax = (S.mbf_fmax - S.mbf_finf) * S.mbf_xmax * logf(S.mbf_xmax) + S.mbf_finf * S.mbf_xmax;
		e_fene = (S.mbf_fmax - S.mbf_finf) * S.mbf_xmax * logf(fabsf(rbackr0)) + S.mbf_finf * fabsf(rbackr0) - long_xmax + fene_xmax;
		force = rback * (-copysignf(1.f, rbackr0) * ((S.mbf_fmax - S.mbf_finf) * S.mbf_xmax / fabsf(rbackr0) + S.mbf_finf) / rbackmod);
	}
	else {
		if(fabsf(rbackr0) >= FENE_DELTA) {
			err |= 1;
			rbackr0 = copysignf(FENE_DELTA * 0.999f, rbackr0);
		}
		e_fene = -(FENE_EPS * 0.5f) * logf(1.f - SQR(rbackr0) / FENE_DELTA2);
		force = rback * (-(FENE_EPS * rbackr0 / (FENE_DELTA2 - SQR(rbackr0))) / rbackmod);
	}
	o.force = force;
	o.tp = -cross(p.back, force);
	o.tq = cross(q.back, force);
	e3[0] = e_fene;

	// bonded excluded volume: base-base, p.base-q.back, p.back-q.base
	f3 fx;
	float ex = repulsive_lj(g.base, fx, EXCL_S2, EXCL_R2, EXCL_B2, EXCL_RC2);
	o.force += fx;
	o.tp -= cross(p.base, fx);
	o.tq += cross(q.base, fx);
	ex += repulsive_lj(g.back_base, fx, EXCL_S3, EXCL_R3, EXCL_B3, EXCL_RC3);
	o.force += fx;
	o.tp -= cross(p.base, fx);
	o.tq += cross(q.back, fx);
	ex += repulsive_lj(g.base_back, fx, EXCL_S4, EXCL_R4, EXCL_B4, EXCL_RC4);
	o.force += fx;
	o.tp -= cross(p.back, fx);
	o.tq += cross(q.base, fx);
	e3[1] = ex;

	// stacking
	const f3 &a1 = p.a1, &a2 = p.a2, &a3 = p.a3, &b1 = q.a1, &b2 = q.a2, &b3 = q.a3;
	float rbackrefmod = sqrtf(dot(rbackref, rbackref));
	f3 rstack = g.stack;
	float rstackmod = sqrtf(dot(rstack, rstack));
	f3 rstackdir = rstack * (1.f / rstackmod);

	float cost4 = dot(a3, b3);
	float cost5 = dot(a3, rstackdir);
	float cost6 = -dot(b3, rstackdir);
	float inv_rbr = 1.f / rbackrefmod;
	float cosphi1 = dot(a2, rbackref) * inv_rbr;
	float cosphi2 = dot(b2, rbackref) * inv_rbr;

	float vf1 = f1(S, rstackmod, 1, q.type, p.type);
	float f4t4, f4t4D, f4t5, f4t5D, f4t6, f4t6D;
	cf4(cost4, S_STCK_T4, f4t4, f4t4D);
	cf4(-cost5, S_STCK_T5, f4t5, f4t5D);
	cf4(cost6, S_STCK_T5, f4t6, f4t6D);
	float f5phi1 = f5(cosphi1);
	float f5phi2 = f5(cosphi2);
	float es = vf1 * f4t4 * f4t5 * f4t6 * f5phi1 * f5phi2;
	e3[2] = es;

	if(es != 0.f) {
		float vf1D = f1D(S, rstackmod, 1, q.type, p.type);
		float f4t4Dsin = -f4t4D;
		float f4t5Dsin = -f4t5D;
		float f4t6Dsin = -f4t6D;
		float f5phi1D = f5D(cosphi1);
		float f5phi2D = f5D(cosphi2);
		float inv_rs = 1.f / rstackmod;

		f3 sf = -rstackdir * (vf1D * f4t4 * f4t5 * f4t6 * f5phi1 * f5phi2);
		sf += -(a3 - rstackdir * cost5) * (vf1 * f4t4 * f4t5Dsin * f4t6 * f5phi1 * f5phi2 * inv_rs);
		sf += -(b3 + rstackdir * cost6) * (vf1 * f4t4 * f4t5 * f4t6Dsin * f5phi1 * f5phi2 * inv_rs);

		float rbrc = rbackrefmod * rbackrefmod * rbackrefmod;
		float inv_rbrc = 1.f / rbrc;
		float ra2 = dot(rstackdir, a2);
		float ra1 = dot(rstackdir, a1);
		float rb1 = dot(rstackdir, b1);
		float a2b1 = dot(a2, b1);
		float par = rstackmod * ra2 - a2b1 * GAMMA;
		float dcosphi1dr = (SQR(rstackmod) * ra2 - ra2 * SQR(rbackrefmod) - rstackmod * (a2b1 + ra2 * (-ra1 + rb1)) * GAMMA +
				a2b1 * (-ra1 + rb1) * SQR(GAMMA)) * inv_rbrc;
		float dcosphi1dra1 = rstackmod * GAMMA * par * inv_rbrc;
		float dcosphi1dra2 = -rstackmod * inv_rbr;
		float dcosphi1drb1 = -rstackmod * GAMMA * par * inv_rbrc;
		float dcosphi1da1b1 = -SQR(GAMMA) * par * inv_rbrc;
		float dcosphi1da2b1 = GAMMA * inv_rbr;
		float fpp1 = -vf1 * f4t4 * f4t5 * f4t6 * f5phi1D * f5phi2;
		sf += -(rstackdir * dcosphi1dr + ((a2 - rstackdir * ra2) * dcosphi1dra2 + (a1 - rstackdir * ra1) * dcosphi1dra1 +
				(b1 - rstackdir * rb1) * dcosphi1drb1) * inv_rs) * fpp1;

		float qa2 = dot(rstackdir, b2);
		float qa1 = dot(rstackdir, b1);
		float qb1 = dot(rstackdir, a1);
		float qa2b1 = dot(b2, a1);
		float par2 = rstackmod * qa2 + qa2b1 * GAMMA;
		float dcosphi2dr = (par2 * (rstackmod + (qb1 - qa1) * GAMMA) - qa2 * SQR(rbackrefmod)) * inv_rbrc;
		float dcosphi2dra1 = -rstackmod * GAMMA * par2 * inv_rbrc;
		float dcosphi2dra2 = -rstackmod * inv_rbr;
		float dcosphi2drb1 = rstackmod * GAMMA * par2 * inv_rbrc;
		float dcosphi2da1b1 = -SQR(GAMMA) * par2 * inv_rbrc;
		float dcosphi2da2b1 = -GAMMA * inv_rbr;
		float fpp2 = -vf1 * f4t4 * f4t5 * f4t6 * f5phi1 * f5phi2D;
		sf += -(rstackdir * dcosphi2dr + ((b2 - rstackdir * qa2) * dcosphi2dra2 + (b1 - rstackdir * qa1) * dcosphi2dra1 +
				(a1 - rstackdir * qb1) * dcosphi2drb1) * inv_rs) * fpp2;

		o.force += sf;
		f3 tp = -cross(p.stack, sf);
		f3 tq = cross(q.stack, sf);

		f3 t4dir = cross(b3, a3);
		float tm = vf1 * f4t4Dsin * f4t5 * f4t6 * f5phi1 * f5phi2;
		tp -= t4dir * tm;
		tq += t4dir * tm;
		tm = -vf1 * f4t4 * f4t5Dsin * f4t6 * f5phi1 * f5phi2;
		tp -= cross(rstackdir, a3) * tm;
		tm = vf1 * f4t4 * f4t5 * f4t6Dsin * f5phi1 * f5phi2;
		tq += cross(rstackdir, b3) * tm;

		tp += cross(rstackdir, a2) * (fpp1 * dcosphi1dra2) + cross(rstackdir, a1) * (fpp1 * dcosphi1dra1);
		tq += cross(rstackdir, b1) * (fpp1 * dcosphi1drb1);
		f3 pure = cross(a2, b1) * (fpp1 * dcosphi1da2b1) + cross(a1, b1) * (fpp1 * dcosphi1da1b1);
		tp -= pure;
		tq += pure;

		tp += cross(rstackdir, a1) * (fpp2 * dcosphi2drb1);
		tq += cross(rstackdir, b2) * (fpp2 * dcosphi2dra2) + cross(rstackdir, b1) * (fpp2 * dcosphi2dra1);
		pure = cross(a1, b2) * (fpp2 * dcosphi2da2b1) + cross(a1, b1) * (fpp2 * dcosphi2da1b1);
		tp -= pure;
		tq += pure;

		o.tp += tp;
		o.tq += tq;
	}
	o.e = e3[0] + e3[1] + e3[2];
	return o;
}

// ----------------------------------------------------- non-bonded, close
// excluded volume, hydrogen bonding, cross stacking and oxDNA2 coaxial
// stacking of a non-bonded pair whose centres are closer than R_CLOSE
// (src/Interactions/DNAInteraction.cpp:707-1021,
//  src/Interactions/DNA2Interaction.cpp:214-306).  Adds the force on q to
// fq and the lab torque on p to tp; energies into e4 (excl, HB, CRST, CXST).
__device__ void close_pair(const oxh_sys &S, const Frame &p, const Frame &q, const PairGeom &g, f3 &fq, f3 &tp, float e4[4]) {
	// excluded volume: base-base, p.back-q.base, p.base-q.back, back-back
	f3 fx;
	float ex = repulsive_lj(g.base, fx, EXCL_S2, EXCL_R2, EXCL_B2, EXCL_RC2);
	fq += fx;
	tp -= cross(p.base, fx);
	ex += repulsive_lj(g.base_back, fx, EXCL_S4, EXCL_R4, EXCL_B4, EXCL_RC4);
	fq += fx;
	tp -= cross(p.back, fx);
	ex += repulsive_lj(g.back_base, fx, EXCL_S3, EXCL_R3, EXCL_B3, EXCL_RC3);
	fq += fx;
	tp -= cross(p.base, fx);
	ex += repulsive_lj(g.back, fx, EXCL_S1, EXCL_R1, EXCL_B1, EXCL_RC1);
	fq += fx;
	tp -= cross(p.back, fx);
	e4[0] = ex;
	e4[1] = e4[2] = e4[3] = 0.f;

	const f3 &a1 = p.a1, &a3 = p.a3, &b1 = q.a1, &b3 = q.a3;
	f3 rhydro = g.base;
	float rh2 = dot(rhydro, rhydro);

	// hydrogen bonding
	if(p.btype + q.btype == 3 && SQR(HYDR_RCLOW) < rh2 && rh2 < SQR(HYDR_RCHIGH)) {
		float hb_multi = (abs(q.btype) >= 300 && abs(p.btype) >= 300) ? S.hb_multi : 1.f;
		float rmod = sqrtf(rh2);
		float inv_r = 1.f / rmod;
		f3 dir = rhydro * inv_r;
		float cost1 = -dot(a1, b1), cost2 = -dot(b1, dir), cost3 = dot(a1, dir);
		float cost4 = dot(a3, b3), cost7 = -dot(b3, dir), cost8 = dot(a3, dir);
		float vf1 = hb_multi * f1(S, rmod, 0, q.type, p.type);
		float f4t1, f4t2, f4t3, f4t4, f4t7, f4t8, d1, d2, d3, d4, d7, d8;
		cf4(cost1, S_HYDR_T1, f4t1, d1);
		cf4(cost2, S_HYDR_T2, f4t2, d2);
		cf4(cost3, S_HYDR_T2, f4t3, d3);
		cf4(cost4, S_HYDR_T4, f4t4, d4);
		cf4(cost7, S_HYDR_T7, f4t7, d7);
		cf4(cost8, S_HYDR_T7, f4t8, d8);
		float e = vf1 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8;
		e4[1] = e;
		if(e != 0.f) {
			float vf1D = hb_multi * f1D(S, rmod, 0, q.type, p.type);
			float f4t1Dsin = d1, f4t2Dsin = d2, f4t3Dsin = -d3, f4t4Dsin = -d4, f4t7Dsin = d7, f4t8Dsin = -d8;
			f3 force = -dir * (vf1D * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8);
			f3 t = {0, 0, 0};
			t -= cross(a3, b3) * (-vf1 * f4t1 * f4t2 * f4t3 * f4t4Dsin * f4t7 * f4t8);
			t -= cross(a1, b1) * (-vf1 * f4t1Dsin * f4t2 * f4t3 * f4t4 * f4t7 * f4t8);
			force += (b1 + dir * cost2) * (vf1 * f4t1 * f4t2Dsin * f4t3 * f4t4 * f4t7 * f4t8 * inv_r);
			force += (a1 - dir * cost3) * (vf1 * f4t1 * f4t2 * f4t3Dsin * f4t4 * f4t7 * f4t8 * inv_r);
			t += cross(dir, a1) * (-vf1 * f4t1 * f4t2 * f4t3Dsin * f4t4 * f4t7 * f4t8);
			force += (b3 + dir * cost7) * (vf1 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7Dsin * f4t8 * inv_r);
			force += (a3 - dir * cost8) * (vf1 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8Dsin * inv_r);
			t += cross(dir, a3) * (-vf1 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8Dsin);
			t -= cross(p.base, force);
			fq += force;
			tp += t;
		}
	}

	// cross stacking
	if(SQR(CRST_RCLOW) < rh2 && rh2 < SQR(CRST_RCHIGH)) {
		float rmod = sqrtf(rh2);
		float inv_r = 1.f / rmod;
		f3 dir = rhydro * inv_r;
		float cost1 = -dot(a1, b1), cost2 = -dot(b1, dir), cost3 = dot(a1, dir);
		float cost4 = dot(a3, b3), cost7 = -dot(b3, dir), cost8 = dot(a3, dir);
		float vf2 = f2(rmod, 0);
		float f4t1, f4t2, f4t3, d1, d2, d3, ga, gaD, gb, gbD;
		cf4(cost1, S_CRST_T1, f4t1, d1);
		cf4(cost2, S_CRST_T2, f4t2, d2);
		cf4(cost3, S_CRST_T2, f4t3, d3);
		cf4(cost4, S_CRST_T4, ga, gaD);
		cf4(-cost4, S_CRST_T4, gb, gbD);
		float f4t4 = ga + gb, f4t4Dsin = -gaD + gbD;
		cf4(cost7, S_CRST_T7, ga, gaD);
		cf4(-cost7, S_CRST_T7, gb, gbD);
		float f4t7 = ga + gb, f4t7Dsin = gaD - gbD;
		cf4(cost8, S_CRST_T7, ga, gaD);
		cf4(-cost8, S_CRST_T7, gb, gbD);
		float f4t8 = ga + gb, f4t8Dsin = -gaD + gbD;
		float e = vf2 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8;
		e4[2] = e;
		if(e != 0.f) {
			float vf2D = f2D(rmod, 0);
			float f4t1Dsin = d1, f4t2Dsin = d2, f4t3Dsin = -d3;
			f3 force = -dir * (vf2D * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8);
			f3 t = {0, 0, 0};
			t -= cross(a1, b1) * (-vf2 * f4t1Dsin * f4t2 * f4t3 * f4t4 * f4t7 * f4t8);
			force += (b1 + dir * cost2) * (vf2 * f4t1 * f4t2Dsin * f4t3 * f4t4 * f4t7 * f4t8 * inv_r);
			force += (a1 - dir * cost3) * (vf2 * f4t1 * f4t2 * f4t3Dsin * f4t4 * f4t7 * f4t8 * inv_r);
			t += cross(dir, a1) * (-vf2 * f4t1 * f4t2 * f4t3Dsin * f4t4 * f4t7 * f4t8);
			t -= cross(a3, b3) * (-vf2 * f4t1 * f4t2 * f4t3 * f4t4Dsin * f4t7 * f4t8);
			force += (b3 + dir * cost7) * (vf2 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7Dsin * f4t8 * inv_r);
			force += (a3 - dir * cost8) * (vf2 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8Dsin * inv_r);
			t += cross(dir, a3) * (-vf2 * f4t1 * f4t2 * f4t3 * f4t4 * f4t7 * f4t8Dsin);
			t -= cross(p.base, force);
			fq += force;
			tp += t;
		}
	}

	// coaxial stacking (oxDNA2 form)
	f3 rstack = g.stack;
	float rs2 = dot(rstack, rstack);
	if(SQR(CXST_RCLOW) < rs2 && rs2 < SQR(CXST_RCHIGH)) {
		float rmod = sqrtf(rs2);
		float inv_r = 1.f / rmod;
		f3 dir = rstack * inv_r;
		float cost1 = -dot(a1, b1), cost4 = dot(a3, b3), cost5 = dot(a3, dir), cost6 = -dot(b3, dir);
		float vf2 = f2(rmod, 1);
		float f4t1, d1, f4t4, d4, ga, gaD, gb, gbD;
		cf4(cost1, S_CXST_T1, f4t1, d1);
		cf4(cost4, S_CXST_T4, f4t4, d4);
		cf4(cost5, S_CXST_T5, ga, gaD);
		cf4(-cost5, S_CXST_T5, gb, gbD);
		float f4t5 = ga + gb, f4t5Dsin = -gaD + gbD;
		cf4(cost6, S_CXST_T5, ga, gaD);
		cf4(-cost6, S_CXST_T5, gb, gbD);
		float f4t6 = ga + gb, f4t6Dsin = gaD - gbD;
		float e = vf2 * f4t1 * f4t4 * f4t5 * f4t6;
		e4[3] = e;
		if(e != 0.f) {
			float vf2D = f2D(rmod, 1);
			float f4t1Dsin = d1, f4t4Dsin = -d4;
			f3 force = -dir * (vf2D * f4t1 * f4t4 * f4t5 * f4t6);
			f3 t = {0, 0, 0};
			t -= cross(a1, b1) * (-vf2 * f4t1Dsin * f4t4 * f4t5 * f4t6);
			t -= cross(a3, b3) * (-vf2 * f4t1 * f4t4Dsin * f4t5 * f4t6);
			float fact = vf2 * f4t1 * f4t4 * f4t5Dsin * f4t6;
			force += (a3 - dir * cost5) * (fact * inv_r);
			t -= cross(dir, a3) * fact;
			fact = vf2 * f4t1 * f4t4 * f4t5 * f4t6Dsin;
			force += (b3 + dir * cost6) * (fact * inv_r);
			t -= cross(p.stack, force);
			fq += force;
			tp += t;
		}
	}
}

// Debye-Hueckel between backbone sites (src/Interactions/DNA2Interaction.cpp:
// 157-212); cf = product of the two charge factors
__device__ __forceinline__ float debye_huckel(const oxh_sys &S, f3 rback, float cf, f3 &fq) {
	float rb2 = dot(rback, rback);
	if(rb2 >= S.dh_rc * S.dh_rc) return 0.f;
	float rbm = sqrtf(rb2);
	float inv = 1.f / rbm;
	float e;
	if(rbm < S.dh_rhigh) {
		float ex = expf(S.dh_minus_kappa * rbm);
		e = ex * (S.dh_prefactor * inv);
		fq = rback * (inv * (-S.dh_prefactor * ex) * (S.dh_minus_kappa * inv - inv * inv) * cf);
	}
	else {
		e = S.dh_b * SQR(rbm - S.dh_rc);
		fq = rback * (inv * (-2.f * S.dh_b * (rbm - S.dh_rc)) * cf);
	}
	return e * cf;
}

// FP32 nearest image of two wrapped backbone sites (far pairs only)
__device__ __forceinline__ f3 min_image_f(float4 a, float4 b, const oxh_sys &S) {
	float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
	dx -= S.boxf[0] * rintf(dx * S.inv_boxf[0]);
	dy -= S.boxf[1] * rintf(dy * S.inv_boxf[1]);
	dz -= S.boxf[2] * rintf(dz * S.inv_boxf[2]);
	return mk(dx, dy, dz);
}

// pair virial r_pq (x) F_q (= -r_pq (x) F_p) in the reference's component
// order xx, yy, zz, xy, xz, yz (src/Interactions/BaseInteraction.cpp:159-200)
__device__ __forceinline__ void add_virial(float sv[6], f3 r, f3 fq) {
	sv[0] += r.x * fq.x;
	sv[1] += r.y * fq.y;
	sv[2] += r.z * fq.z;
	sv[3] += r.x * fq.y;
	sv[4] += r.x * fq.z;
	sv[5] += r.y * fq.z;
}

// sum over the K lanes of a group (K a power of two, groups aligned to K
// inside the wave), fixed butterfly order
__device__ __forceinline__ float group_sum(float v, int K) {
#pragma unroll
	for(int off = K / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, K);
	return v;
}

// Forces, torques and energy of particle i from the K threads l = 0..K-1
// of its group (consecutive lanes of one wave, K | 64): lane 0 takes the
// 3' bond, lane 1 the 5' bond, neighbour k goes to lane k mod K; the lanes'
// partial sums are combined by a fixed butterfly, so the result does not
// depend on timing.  Lane 0 writes F (xyz, .w = energy) and the body-frame
// torque (.w = HB energy).  Far pairs are added here when FAR, else by
// k_forces_far.  Every lane of the group must call it.
// K may be a run-time value (the workgroup-resident engine) or a constant
// (the grid kernels, where it folds into the loop strides and shuffles).
template<bool SPLIT, bool FAR>
__device__ __forceinline__ void particle_forces(int K, int i, int l, int N, const double4 *__restrict__ pos,
		const double4 *__restrict__ q64, const float4 *__restrict__ q32,
		const float4 *__restrict__ bpos, const int2 *__restrict__ bonds, const int2 *__restrict__ pinfo, const int *__restrict__ nbr,
		const int2 *__restrict__ nnbr, int max_neigh, const oxh_sys *__restrict__ systems, oxh_sys_state *__restrict__ sstate,
		float4 *__restrict__ Fout, float4 *__restrict__ Tout, float *__restrict__ esplit, int *__restrict__ hbcount,
		float *__restrict__ stress) {
	const int2 pi = pinfo[i];
	const oxh_sys &S = systems[pi.y];
	const double4 P = pos[i];
	Sites64 sp;
	const Frame fp = make_frame_d(q64[i], pi.x, sp);
	const int2 bd = bonds[i];

	f3 F = {0, 0, 0}, T = {0, 0, 0};
	float E = 0.f, Ehb = 0.f;
	float es[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	float sv[6] = {0, 0, 0, 0, 0, 0};   // pair virial, r_pq (x) F_q summed over this particle's pairs
	int err = 0;
	int nhb = 0;

	// bonded neighbours: i is the 5' partner of n3 and the 3' partner of n5
	if(l == 0 && bd.x >= 0) {
		int j = bd.x;
		Sites64 sq;
		Frame fq = make_frame_d(q64[j], pinfo[j].x, sq);
		const d3 rd = min_image_d(P, pos[j], S);
		const PairGeom g = pair_geom(rd, sp, sq);
		float stretch;
		f3 backref;
		bonded_geom(rd, sp, sq, S, stretch, backref);
		float e3[3];
		PairOut o = bonded_pair(S, fp, fq, g, stretch, backref, e3, err);
		F -= o.force;
		T += o.tp;
		E += 0.5f * o.e;
		if(SPLIT) {
			es[0] += 0.5f * e3[0];
			es[1] += 0.5f * e3[1];
			es[2] += 0.5f * e3[2];
			add_virial(sv, g.r, o.force);
		}
	}
	if(l == (K > 1 ? 1 : 0) && bd.y >= 0) {
		int j = bd.y;
		Sites64 sq;
		Frame fq = make_frame_d(q64[j], pinfo[j].x, sq);
		const d3 rd = min_image_d(pos[j], P, S);
		const PairGeom g = pair_geom(rd, sq, sp);
		float stretch;
		f3 backref;
		bonded_geom(rd, sq, sp, S, stretch, backref);
		float e3[3];
		PairOut o = bonded_pair(S, fq, fp, g, stretch, backref, e3, err);
		F += o.force;
		T += o.tq;
		E += 0.5f * o.e;
		if(SPLIT) {
			es[0] += 0.5f * e3[0];
			es[1] += 0.5f * e3[1];
			es[2] += 0.5f * e3[2];
			add_virial(sv, g.r, o.force);
		}
	}

	const int2 nn = nnbr[i];
	const int *row = nbr + (size_t) i * max_neigh;
	const float cf_i = bpos[i].w;

	// close sublist: every term
	for(int k = l; k < nn.x; k += K) {
		const int j = row[k];
		Sites64 sq;
		const Frame fq = make_frame_d(q64[j], pinfo[j].x, sq);
		const PairGeom g = pair_geom(min_image_d(P, pos[j], S), sp, sq);
		const f3 r = g.r;
		f3 fq_ = {0, 0, 0};
		float edh = debye_huckel(S, g.back, cf_i * bpos[j].w, fq_);
		f3 tp = -cross(fp.back, fq_);
		float e = edh;
		if(SPLIT) es[7] += 0.5f * edh;
		if(dot(r, r) < R_CLOSE * R_CLOSE) {
			float e4[4];
			close_pair(S, fp, fq, g, fq_, tp, e4);
			e += e4[0] + e4[1] + e4[2] + e4[3];
			Ehb += 0.5f * e4[1];
			if(hbcount && e4[1] < -0.1f) nhb++;
			if(SPLIT) {
				es[3] += 0.5f * e4[0];
				es[4] += 0.5f * e4[1];
				es[5] += 0.5f * e4[2];
				es[6] += 0.5f * e4[3];
			}
		}
		F -= fq_;
		T += tp;
		E += 0.5f * e;
		if(SPLIT) add_virial(sv, r, fq_);
	}

	// far sublist: Debye-Hueckel between backbone sites only (or left to
	// k_forces_far when FAR is false)
	const float4 Bi = bpos[i];
	f3 Ff = {0, 0, 0};
	float Ef = 0.f;
	for(int k = l; FAR && k < nn.y; k += K) {
		const int j = row[max_neigh - 1 - k];
		const float4 Bj = bpos[j];
		f3 rb;
		if(!S.small_box) rb = min_image_f(Bi, Bj, S);
		else {
			// the reference's image choice: nearest image of the centres,
			// then the site offsets (src/Interactions/DNA2Interaction.cpp:169)
			const Frame fq = make_frame(q32[j], 0);
			rb = min_image(P, pos[j], S) + fq.back - fp.back;
		}
		f3 fq_ = {0, 0, 0};
		float edh = debye_huckel(S, rb, cf_i * Bj.w, fq_);
		Ff += fq_;
		Ef += edh;
		if(SPLIT) add_virial(sv, min_image(P, pos[j], S), fq_);
	}
	F -= Ff;
	T -= cross(fp.back, Ff);
	E += 0.5f * Ef;
	if(SPLIT) es[7] += 0.5f * Ef;

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
		if(hbcount) nhb = (int) group_sum((float) nhb, K);
		err = (int) group_sum((float) err, K);
	}
	if(l != 0) return;

	// torque into the body frame: orientationT * T_lab
	f3 Tb = {dot(fp.a1, T), dot(fp.a2, T), dot(fp.a3, T)};
	Fout[i] = make_float4(F.x, F.y, F.z, E);
	Tout[i] = make_float4(Tb.x, Tb.y, Tb.z, Ehb);
	if(SPLIT) {
		for(int q = 0; q < 8; q++) esplit[(size_t) q * N + i] = es[q];
		// each particle takes half of every pair's virial
		for(int q = 0; q < 6; q++) stress[(size_t) q * N + i] = 0.5f * sv[q];
	}
	if(hbcount) hbcount[i] = nhb;
	if(err) atomicOr(&sstate[pi.y].error, 1);
}

}  // namespace oxh
