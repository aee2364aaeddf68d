// C ABI of the oxDNA2 hot path (include/oxh.h): batch set-up, stepping,
// downloads.  Host code; kernels live in oxh_forces.hip, oxh_lists.hip and
// oxh_integrate.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/oxh.h"
#include "oxh_device.h"
#include "oxh_internal.h"
#include "oxh_rna.h"
#include "oxh_step.h"

extern "C" int oxh_internal_hilbert_probe(int n, const void *d_xyzb, void *d_out, hipStream_t st);

struct oxh_batch {
	oxh::Batch b;
	// per system host copies of what downloads need
	std::vector<int> first, n;
	// backup of the state at the start of a chunk (overflow recovery)
	double4 *bk_pos = nullptr, *bk_vel = nullptr, *bk_Lm = nullptr, *bk_q64 = nullptr;
	float4 *bk_F = nullptr, *bk_T = nullptr;
	double4 *bk_F64 = nullptr, *bk_T64 = nullptr;
	long long *bk_pstep = nullptr;
	oxh_sys_state *bk_state = nullptr;
	size_t device_bytes = 0;
	bool bussi_aligned = true;
	int bussi_ns = 0;
	bool any_bussi = false;
	int chunk = 0;
	// HIP graph replay of whole chunks (OXH_GRAPH=0 launches the same
	// kernels one by one, e.g. under a profiler that does not follow graphs)
	bool use_graph = true;
	// Hilbert re-sort every sort_interval steps (0 = never), checked at
	// chunk boundaries
	long long sort_interval = 0, since_sort = 0;
	// workgroup-resident engine (every system small and without cells;
	// OXH_PERSIST=0 selects the one-kernel-per-phase path instead), steps
	// per launch
	bool persist = false;
	int persist_chunk = 256;
	int persist_block = 512;   // threads per workgroup (OXH_PERSIST_BLOCK = 256 or 512)
	// slot pool (oxh_pool_create): capacity and reserved cells per slot,
	// step target and steps done per slot, per-chunk step limits
	bool pool = false;
	std::vector<int> cap, cell_budget, h_limit;
	// strands (chains of bonds) of each system, by original particle index;
	// fix_diffusion_every: interval of the strand re-boxing (0 = off)
	std::vector<std::vector<int>> strand_of;
	std::vector<int> n_strands;
	long long fix_every = 0, since_fix = 0;
	std::vector<long long> target, done_steps;
	double4 *bk_rlist = nullptr;
	float4 *bk_blist = nullptr;
	// external forces of each system: (particle, record) in the caller's
	// order
	std::vector<std::vector<std::pair<int, oxh_ext_rec>>> ext_of;
	// centre-of-mass groups of each system's "com" forces: (com list,
	// reference list), indices in the system; a com record's ref is its
	// group's position in this list
	std::vector<std::vector<std::pair<std::vector<int>, std::vector<int>>>> grp_of;
};

namespace {

thread_local std::string g_err;

oxh_status fail(oxh_status code, const char *fmt, ...) {
	char buf[1024];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	g_err = buf;
	return code;
}

#define HIPCHK(call)                                                                                      \
	do {                                                                                                  \
		hipError_t e_ = (call);                                                                           \
		if(e_ != hipSuccess) return fail(OXH_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
	} while(0)

template<typename T>
oxh_status dalloc(oxh_batch *B, T **p, size_t count) {
	size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
	hipError_t e = hipMalloc((void **) p, bytes);
	if(e != hipSuccess) return fail(OXH_ERR_ALLOC, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
	B->device_bytes += bytes;
	return OXH_OK;
}

#define TRY(x)                          \
	do {                                \
		oxh_status s_ = (x);            \
		if(s_ != OXH_OK) return s_;     \
	} while(0)

inline double SQRd(double x) { return x * x; }

// Model constants of one system, in FP64 then rounded to FP32 for the
// kernels.  The formulas are the reference's: DNAInteraction::init
// (src/Interactions/DNAInteraction.cpp:298-377), DNA2Interaction::
// get_settings / init (src/Interactions/DNA2Interaction.cpp:62-155).
// The reference's float-literal constants go through FP64 library calls.
void rna_model_constants(const oxh_system_desc &d, oxh_sys &S, double &rcut);

void model_constants(const oxh_system_desc &d, oxh_sys &S, double &rcut) {
	if(d.interaction == OXH_INTERACTION_RNA2 || d.interaction == OXH_INTERACTION_RNA) {
		rna_model_constants(d, S, rcut);
		return;
	}
	const double T = d.T;
	const float STCK_RC = 0.9f, STCK_R0 = 0.4f, STCK_A = 6.f, HYDR_RC = 0.75f, HYDR_R0 = 0.4f, HYDR_A = 8.f;
	const double stck_shift = SQRd(1 - std::exp((double) (-(STCK_RC - STCK_R0) * STCK_A)));
	const double hydr_shift = SQRd(1 - std::exp((double) (-(HYDR_RC - HYDR_R0) * HYDR_A)));
	double eps[2][5][5];
	for(int i = 0; i < 5; i++)
		for(int j = 0; j < 5; j++) {
			eps[1][i][j] = (double) 1.3523f + (double) 2.6717f * T;
			eps[0][i][j] = (double) 1.0678f;
		}
	if(d.seq_stck) {
		for(int i = 0; i < 4; i++)
			for(int j = 0; j < 4; j++)
				eps[1][i][j] = d.seq_stck[4 * i + j] * (1.0 - d.seq_stck_fact_eps + (T * 9.0 * d.seq_stck_fact_eps));
		eps[0][0][3] = eps[0][3][0] = d.seq_hydr_AT;
		eps[0][1][2] = eps[0][2][1] = d.seq_hydr_GC;
	}
	for(int t = 0; t < 2; t++)
		for(int i = 0; i < 5; i++)
			for(int j = 0; j < 5; j++) {
				S.f1_eps_d[t][i][j] = eps[t][i][j];
				S.f1_shift_d[t][i][j] = eps[t][i][j] * (t == 1 ? stck_shift : hydr_shift);
				S.f1_eps[t][i][j] = (float) S.f1_eps_d[t][i][j];
				S.f1_shift[t][i][j] = (float) S.f1_shift_d[t][i][j];
			}

	const float salt_f = (float) d.salt;
	const double lf = d.dh_lambda >= 0 ? d.dh_lambda : 0.3616455;
	const double q = d.dh_strength >= 0 ? d.dh_strength : 0.0543;
	const double lambda_set = lf * std::sqrt(T / 0.1f) / std::sqrt((double) salt_f);
	const double rhigh = d.dh_rhigh > 0 ? d.dh_rhigh : 3.0 * lambda_set;
	const double l = lf * std::sqrt(T / 0.1) / std::sqrt((double) salt_f);
	const double x = rhigh;
	const double B = -(std::exp(-x / l) * q * q * (x + l) * (x + l)) / (-4. * x * x * x * l * l * q);
	const double RC = x * (q * x + 3. * q * l) / (q * (x + l));
	S.dh_prefactor_d = q;
	S.dh_minus_kappa_d = -1.0 / l;
	S.dh_rhigh_d = rhigh;
	S.dh_rc_d = RC;
	S.dh_b_d = B;
	S.dh_half = d.dh_half_charged_ends ? 1 : 0;
	S.hb_multi_d = d.hb_multiplier;
	S.fene_r0_d = (double) 0.7564f;
	S.use_mbf = 0;
	if(d.max_backbone_force >= 0) {
		const double FE = 2.0, FD2 = (double) 0.0625f;
		double fmax = d.max_backbone_force;
		S.use_mbf = 1;
		S.mbf_fmax_d = fmax;
		S.mbf_xmax_d = (-FE + std::sqrt(FE * FE + 4.f * fmax * fmax * FD2)) / (2.f * fmax);
		S.mbf_finf_d = d.max_backbone_force_far >= 0 ? d.max_backbone_force_far : (double) 0.04f;
	}
	S.dh_prefactor = (float) S.dh_prefactor_d;
	S.dh_minus_kappa = (float) S.dh_minus_kappa_d;
	S.dh_rhigh = (float) S.dh_rhigh_d;
	S.dh_rc = (float) S.dh_rc_d;
	S.dh_b = (float) S.dh_b_d;
	S.hb_multi = (float) S.hb_multi_d;
	S.fene_r0 = (float) S.fene_r0_d;
	S.mbf_fmax = (float) S.mbf_fmax_d;
	S.mbf_xmax = (float) S.mbf_xmax_d;
	S.mbf_finf = (float) S.mbf_finf_d;
	const double arm = std::sqrt(SQRd((double) -0.34f) + SQRd((double) 0.3408f));
	rcut = std::max(2 * arm + (double) 0.711879214356f, 2 * (double) 0.4f + (double) 0.783775f);
	rcut = std::max(rcut, 2.0 * arm + RC);
}

// The oxRNA constants of one system into the same per-system fields:
// RNAInteraction::init (src/Interactions/RNAInteraction.cpp:110-118,
// 275-291: stacking and HB well depths, r_cut), RNAInteraction::
// get_settings (:91-103: max_backbone_force) and, for oxRNA2, RNA2Interaction::
// get_settings / init (src/Interactions/RNAInteraction2.cpp:33-102:
// Debye-Hueckel with V = 0).  Plain oxRNA has no Debye-Hueckel term
// (dh_rc = 0).
void rna_model_constants(const oxh_system_desc &d, oxh_sys &S, double &rcut) {
	namespace R = oxh::rna;
	const double T = d.T;
	const double stck_shift = SQRd(1 - std::exp((double) (-(R::F1_RC[1] - R::F1_R0[1]) * R::F1_A[1])));
	const double hydr_shift = SQRd(1 - std::exp((double) (-(R::F1_RC[0] - R::F1_R0[0]) * R::F1_A[0])));
	for(int i = 0; i < 5; i++)
		for(int j = 0; j < 5; j++) {
			S.f1_eps_d[1][i][j] = (double) R::STCK_BASE_EPS + (double) R::STCK_FACT_EPS * T;
			S.f1_eps_d[0][i][j] = (double) R::HYDR_EPS;
			S.f1_shift_d[1][i][j] = S.f1_eps_d[1][i][j] * stck_shift;
			S.f1_shift_d[0][i][j] = S.f1_eps_d[0][i][j] * hydr_shift;
			for(int t = 0; t < 2; t++) {
				S.f1_eps[t][i][j] = (float) S.f1_eps_d[t][i][j];
				S.f1_shift[t][i][j] = (float) S.f1_shift_d[t][i][j];
			}
		}
	const double arm = std::sqrt(SQRd((double) R::POS_BACK_A1) + SQRd((double) R::POS_BACK_A2) + SQRd((double) R::POS_BACK_A3));
	rcut = std::max(2 * arm + (double) R::EXCL_RC1, 2 * std::fabs((double) R::POS_BASE) + std::max((double) R::HYDR_RCHIGH, (double) R::F2_RCHIGH[0]));
	S.dh_half = d.dh_half_charged_ends ? 1 : 0;
	if(d.interaction == OXH_INTERACTION_RNA2) {
		const float salt_f = (float) d.salt;
		const double lf = d.dh_lambda >= 0 ? d.dh_lambda : 0.3667258;
		const double q = d.dh_strength >= 0 ? d.dh_strength : 0.0858;
		const double l = lf * std::sqrt(T / 0.1f) / std::sqrt((double) salt_f);
		const double x = d.dh_rhigh > 0 ? d.dh_rhigh : 3.0 * l;
		S.dh_prefactor_d = q;
		S.dh_minus_kappa_d = -1.0 / l;
		S.dh_rhigh_d = x;
		S.dh_b_d = -(std::exp(-x / l) * q * q * (x + l) * (x + l)) / (4. * x * x * x * l * l * (-q));
		S.dh_rc_d = x * (q * x + 3. * q * l) / (q * (x + l));
		rcut = std::max(rcut, 2. * arm + S.dh_rc_d);
	}
	S.fene_r0_d = (double) R::FENE_R0;
	S.use_mbf = 0;
	if(d.max_backbone_force >= 0) {
		const double FE = (double) R::FENE_EPS, FD2 = (double) R::FENE_DELTA2;
		const double fmax = d.max_backbone_force;
		S.use_mbf = 1;
		S.mbf_fmax_d = fmax;
		S.mbf_xmax_d = (-FE + std::sqrt(FE * FE + 4.f * fmax * fmax * FD2)) / (2.f * fmax);
		S.mbf_finf_d = d.max_backbone_force_far >= 0 ? d.max_backbone_force_far : (double) 0.04f;
	}
	S.hb_multi_d = 1.0;
	S.dh_prefactor = (float) S.dh_prefactor_d;
	S.dh_minus_kappa = (float) S.dh_minus_kappa_d;
	S.dh_rhigh = (float) S.dh_rhigh_d;
	S.dh_rc = (float) S.dh_rc_d;
	S.dh_b = (float) S.dh_b_d;
	S.hb_multi = 1.f;
	S.fene_r0 = (float) S.fene_r0_d;
	S.mbf_fmax = (float) S.mbf_fmax_d;
	S.mbf_xmax = (float) S.mbf_xmax_d;
	S.mbf_finf = (float) S.mbf_finf_d;
}

// orientation (a1, a3) -> unit quaternion (x, y, z, w); M has columns a1, a2, a3
void frame_to_quat(const double *a1, const double *a3, double q[4]) {
	double a2[3] = {a3[1] * a1[2] - a3[2] * a1[1], a3[2] * a1[0] - a3[0] * a1[2], a3[0] * a1[1] - a3[1] * a1[0]};
	double M[3][3] = {{a1[0], a2[0], a3[0]}, {a1[1], a2[1], a3[1]}, {a1[2], a2[2], a3[2]}};
	double tr = M[0][0] + M[1][1] + M[2][2];
	double x, y, z, w;
	if(tr > 0) {
		double s = 0.5 / std::sqrt(tr + 1.0);
		w = 0.25 / s;
		x = (M[2][1] - M[1][2]) * s;
		y = (M[0][2] - M[2][0]) * s;
		z = (M[1][0] - M[0][1]) * s;
	}
	else if(M[0][0] > M[1][1] && M[0][0] > M[2][2]) {
		double s = 0.5 / std::sqrt(1.0 + M[0][0] - M[1][1] - M[2][2]);
		w = (M[2][1] - M[1][2]) * s;
		x = 0.25 / s;
		y = (M[0][1] + M[1][0]) * s;
		z = (M[0][2] + M[2][0]) * s;
	}
	else if(M[1][1] > M[2][2]) {
		double s = 0.5 / std::sqrt(1.0 + M[1][1] - M[0][0] - M[2][2]);
		w = (M[0][2] - M[2][0]) * s;
		x = (M[0][1] + M[1][0]) * s;
		y = 0.25 / s;
		z = (M[1][2] + M[2][1]) * s;
	}
	else {
		double s = 0.5 / std::sqrt(1.0 + M[2][2] - M[0][0] - M[1][1]);
		w = (M[1][0] - M[0][1]) * s;
		x = (M[0][2] + M[2][0]) * s;
		y = (M[1][2] + M[2][1]) * s;
		z = 0.25 / s;
	}
	double n = std::sqrt(x * x + y * y + z * z + w * w);
	q[0] = x / n;
	q[1] = y / n;
	q[2] = z / n;
	q[3] = w / n;
}

void quat_to_axes(const double4 &q, double a1[3], double a3[3]) {
	double sqx = q.x * q.x, sqy = q.y * q.y, sqz = q.z * q.z, sqw = q.w * q.w;
	double xy = q.x * q.y, xz = q.x * q.z, xw = q.x * q.w, yz = q.y * q.z, yw = q.y * q.w, zw = q.z * q.w;
	a1[0] = sqx - sqy - sqz + sqw;
	a1[1] = 2. * (xy + zw);
	a1[2] = 2. * (xz - yw);
	a3[0] = 2. * (xz + yw);
	a3[1] = 2. * (yz - xw);
	a3[2] = -sqx - sqy + sqz + sqw;
}

__global__ void k_clear_need(int nsys, oxh_sys_state *st, int set_value, int *any) {
	int s = blockIdx.x * blockDim.x + threadIdx.x;
	if(s < nsys) st[s].need_rebuild = set_value;
	if(s == 0) *any = set_value;
}

// counts a list build for every system whose rebuild flag is set, then
// clears the flags
__global__ void k_count_rebuilds(int nsys, oxh_sys_state *st, int *any) {
	int s = blockIdx.x * blockDim.x + threadIdx.x;
	if(s < nsys && st[s].need_rebuild) {
		st[s].need_rebuild = 0;
		st[s].n_rebuilds += 1;
	}
	if(s == 0) *any = 0;
}

oxh_status sync_and_check(oxh_batch *B) {
	HIPCHK(hipStreamSynchronize(B->b.stream));
	HIPCHK(hipGetLastError());
	return OXH_OK;
}

oxh_status read_states(oxh_batch *B, std::vector<oxh_sys_state> &st) {
	st.resize(B->b.nsys);
	HIPCHK(hipMemcpyAsync(st.data(), B->b.d_state, sizeof(oxh_sys_state) * B->b.nsys, hipMemcpyDeviceToHost, B->b.stream));
	return sync_and_check(B);
}

// list build of the current state (systems flagged, or all)
void enqueue_lists(oxh_batch *B, bool all) {
	oxh::Batch &b = B->b;
	if(all) k_clear_need<<<(b.nsys + 63) / 64, 64, 0, b.stream>>>(b.nsys, b.d_state, 1, b.any_rebuild);
	oxh::launch_cells(b.stream, b, b.pos);
	oxh::launch_build_list(b.stream, b, 0, b.pos);
}

// Steps 0 .. count - 1 of a chunk; step_index0 = the batch's step count at
// the chunk's start (Bussi alignment).  The second half of a step and the
// first half of the next share one kernel, except around Bussi
// activations, which need all of a system's velocities in between.
void enqueue_steps(oxh_batch *B, int count, long long step_index0) {
	oxh::Batch &b = B->b;
	auto bussi_due = [&](int k) { return B->any_bussi && (!B->bussi_aligned || (step_index0 + k) % B->bussi_ns == 0); };
	oxh::launch_first_step(b.stream, b, 0);   // also zeroes the cell counters
	for(int k = 0; k < count; k++) {
		oxh::launch_cells_fill_sort(b.stream, b);
		oxh::launch_build_list(b.stream, b, 0, b.pos);
		oxh::launch_forces(b.stream, b, false, k);
		if(k + 1 < count && !bussi_due(k)) oxh::launch_second_first(b.stream, b, k);
		else {
			oxh::launch_second_step(b.stream, b, k);
			if(bussi_due(k)) oxh::launch_bussi(b.stream, b, k);
			if(k + 1 < count) oxh::launch_first_step(b.stream, b, k + 1);
		}
	}
}

// forces of the current state with the engine oxh_batch_run uses
void enqueue_forces(oxh_batch *B, bool split = false) {
	if(B->persist) oxh::launch_persist_forces(B->b.stream, B->b, split, B->persist_block);
	else oxh::launch_forces(B->b.stream, B->b, split);
}

__global__ void k_clear_err(int nsys, oxh_sys_state *st, int mask) {
	int s = blockIdx.x * blockDim.x + threadIdx.x;
	if(s < nsys) st[s].error &= ~mask;
}

// (re)allocates the edge rows and slots for rows of width ecap (0 frees
// them and switches the half-list form off)
oxh_status alloc_edges(oxh_batch *B, int ecap) {
	oxh::Batch &b = B->b;
	HIPCHK(hipStreamSynchronize(b.stream));
	const size_t old_rows = (size_t) b.N * b.ecap;
	void *ptrs[] = {b.eown, b.ein, b.E1, b.E2, b.E3};
	for(void *p : ptrs)
		if(p) HIPCHK(hipFree(p));
	if(b.ecap > 0) B->device_bytes -= old_rows * (2 * sizeof(int) + 3 * sizeof(float4));
	b.eown = b.ein = nullptr;
	b.E1 = b.E2 = b.E3 = nullptr;
	b.ecap = 0;
	if(ecap <= 0) {
		b.edge = false;
		return OXH_OK;
	}
	const size_t rows = (size_t) b.N * ecap;
	TRY(dalloc(B, &b.eown, rows));
	TRY(dalloc(B, &b.ein, rows));
	TRY(dalloc(B, &b.E1, rows));
	TRY(dalloc(B, &b.E2, rows));
	TRY(dalloc(B, &b.E3, rows));
	if(!b.ecount) {
		TRY(dalloc(B, &b.ecount, b.N));
		TRY(dalloc(B, &b.B1, b.N));
		TRY(dalloc(B, &b.B2, b.N));
		TRY(dalloc(B, &b.B3, b.N));
	}
	// empty edge rows until the next list build fills them
	HIPCHK(hipMemsetAsync(b.ecount, 0, (size_t) b.N * sizeof(int2), b.stream));
	b.ecap = ecap;
	return OXH_OK;
}

// (re)allocates both cell slot buffers (binning order and sorted) for
// `cap` slots per cell: every kernel indexes both with b.cell_cap
oxh_status alloc_cell_slots(oxh_batch *B, int cap) {
	oxh::Batch &b = B->b;
	HIPCHK(hipStreamSynchronize(b.stream));
	const size_t ncell = (size_t) std::max(b.ncell_total, 1);
	if(b.cell_items) {
		HIPCHK(hipFree(b.cell_items));
		HIPCHK(hipFree(b.cell_raw));
		B->device_bytes -= 2 * ncell * b.cell_cap * sizeof(int);
		b.cell_items = b.cell_raw = nullptr;
	}
	b.cell_cap = cap;
	TRY(dalloc(B, &b.cell_items, ncell * cap));
	TRY(dalloc(B, &b.cell_raw, ncell * cap));
	return OXH_OK;
}

// doubles the neighbour, cell or edge-row capacity after an overflow and
// clears the overflow bits (the caller rebuilds)
oxh_status grow_lists(oxh_batch *B, const std::vector<oxh_sys_state> &st) {
	oxh::Batch &b = B->b;
	bool grow_neigh = false, grow_cells = false, grow_edge = false;
	for(auto &s : st) {
		if(s.error & oxh::ERR_LIST) grow_neigh = true;
		if(s.error & oxh::ERR_CELLS) grow_cells = true;
		// a row overflow empties the row, which its partners report as well
		if((s.error & oxh::ERR_EDGE) && !(s.error & oxh::ERR_LIST)) grow_edge = true;
	}
	HIPCHK(hipStreamSynchronize(b.stream));
	if(grow_neigh) {
		HIPCHK(hipFree(b.nbr));
		B->device_bytes -= (size_t) b.nstride * b.max_neigh * sizeof(int);
		b.max_neigh = std::min(std::max(b.max_neigh * 2, 8), std::max(b.N, 1));
		TRY(dalloc(B, &b.nbr, (size_t) b.nstride * b.max_neigh));
	}
	if(grow_cells && b.ncell_total) TRY(alloc_cell_slots(B, b.cell_cap * 2));
	if(grow_edge && b.ecap > 0) TRY(alloc_edges(B, std::min(2 * b.ecap, b.max_neigh)));
	k_clear_err<<<(b.nsys + 63) / 64, 64, 0, b.stream>>>(b.nsys, b.d_state, oxh::ERR_OVERFLOW);
	if(b.graph_exec) {
		(void) hipGraphExecDestroy(b.graph_exec);
		b.graph_exec = nullptr;
	}
	return OXH_OK;
}

// builds lists (all systems) and forces, growing buffers until nothing
// overflows
oxh_status build_all(oxh_batch *B) {
	for(int attempt = 0; attempt < 12; attempt++) {
		enqueue_lists(B, true);
		enqueue_forces(B);
		k_clear_need<<<(B->b.nsys + 63) / 64, 64, 0, B->b.stream>>>(B->b.nsys, B->b.d_state, 0, B->b.any_rebuild);
		std::vector<oxh_sys_state> st;
		TRY(read_states(B, st));
		bool overflow = false;
		for(auto &s : st) overflow |= (s.error & oxh::ERR_OVERFLOW) != 0;
		if(!overflow) return OXH_OK;
		TRY(grow_lists(B, st));
	}
	return fail(OXH_ERR_SYSTEM, "neighbour list keeps overflowing");
}

oxh_status backup(oxh_batch *B) {
	oxh::Batch &b = B->b;
	size_t N = b.N;
	HIPCHK(hipMemcpyAsync(B->bk_pos, b.pos, N * sizeof(double4), hipMemcpyDeviceToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(B->bk_vel, b.vel, N * sizeof(double4), hipMemcpyDeviceToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(B->bk_Lm, b.Lm, N * sizeof(double4), hipMemcpyDeviceToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(B->bk_q64, b.q64, N * sizeof(double4), hipMemcpyDeviceToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(B->bk_F, b.F, N * sizeof(float4), hipMemcpyDeviceToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(B->bk_T, b.T, N * sizeof(float4), hipMemcpyDeviceToDevice, b.stream));
	if(b.fp64) {
		HIPCHK(hipMemcpyAsync(B->bk_F64, b.F64, N * sizeof(double4), hipMemcpyDeviceToDevice, b.stream));
		HIPCHK(hipMemcpyAsync(B->bk_T64, b.T64, N * sizeof(double4), hipMemcpyDeviceToDevice, b.stream));
	}
	HIPCHK(hipMemcpyAsync(B->bk_pstep, b.pstep, N * sizeof(long long), hipMemcpyDeviceToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(B->bk_state, b.d_state, b.nsys * sizeof(oxh_sys_state), hipMemcpyDeviceToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(B->bk_rlist, b.rlist, N * sizeof(double4), hipMemcpyDeviceToDevice, b.stream));
	if(b.blist) HIPCHK(hipMemcpyAsync(B->bk_blist, b.blist, N * sizeof(float4), hipMemcpyDeviceToDevice, b.stream));
	return OXH_OK;
}

oxh_status restore(oxh_batch *B) {
	oxh::Batch &b = B->b;
	size_t N = b.N;
	HIPCHK(hipMemcpyAsync(b.pos, B->bk_pos, N * sizeof(double4), hipMemcpyDeviceToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(b.vel, B->bk_vel, N * sizeof(double4), hipMemcpyDeviceToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(b.Lm, B->bk_Lm, N * sizeof(double4), hipMemcpyDeviceToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(b.q64, B->bk_q64, N * sizeof(double4), hipMemcpyDeviceToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(b.F, B->bk_F, N * sizeof(float4), hipMemcpyDeviceToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(b.T, B->bk_T, N * sizeof(float4), hipMemcpyDeviceToDevice, b.stream));
	if(b.fp64) {
		HIPCHK(hipMemcpyAsync(b.F64, B->bk_F64, N * sizeof(double4), hipMemcpyDeviceToDevice, b.stream));
		HIPCHK(hipMemcpyAsync(b.T64, B->bk_T64, N * sizeof(double4), hipMemcpyDeviceToDevice, b.stream));
	}
	HIPCHK(hipMemcpyAsync(b.pstep, B->bk_pstep, N * sizeof(long long), hipMemcpyDeviceToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(b.d_state, B->bk_state, b.nsys * sizeof(oxh_sys_state), hipMemcpyDeviceToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(b.rlist, B->bk_rlist, N * sizeof(double4), hipMemcpyDeviceToDevice, b.stream));
	if(b.blist) HIPCHK(hipMemcpyAsync(b.blist, B->bk_blist, N * sizeof(float4), hipMemcpyDeviceToDevice, b.stream));
	oxh::launch_prepare_mirror(b.stream, b);
	return sync_and_check(B);
}

__global__ void k_set_need_one(oxh_sys_state *st, int s, int *any) {
	st[s].need_rebuild = 1;
	*any = 1;
}

// Rebuilds every system's list (or only system `only`, if >= 0) from the
// positions of its last build (rlist): the same pairs in the same order as
// that build produced, so a capacity change or a replay after an overflow
// does not change any trajectory.  Grows buffers (and then rebuilds every
// system) until nothing overflows.
oxh_status rebuild_from_rlist(oxh_batch *B, int only = -1) {
	oxh::Batch &b = B->b;
	for(int attempt = 0; attempt < 12; attempt++) {
		if(only >= 0 && attempt == 0) k_set_need_one<<<1, 1, 0, b.stream>>>(b.d_state, only, b.any_rebuild);
		else k_clear_need<<<(b.nsys + 63) / 64, 64, 0, b.stream>>>(b.nsys, b.d_state, 1, b.any_rebuild);
		oxh::launch_cells(b.stream, b, b.rlist);
		oxh::launch_build_list(b.stream, b, 0, b.rlist);
		k_clear_need<<<(b.nsys + 63) / 64, 64, 0, b.stream>>>(b.nsys, b.d_state, 0, b.any_rebuild);
		std::vector<oxh_sys_state> st;
		TRY(read_states(B, st));
		bool overflow = false;
		for(auto &x : st) overflow |= (x.error & oxh::ERR_OVERFLOW) != 0;
		if(!overflow) return OXH_OK;
		TRY(grow_lists(B, st));
	}
	return fail(OXH_ERR_SYSTEM, "neighbour list keeps overflowing");
}


// Per-system parameters of desc d (validation, T-dependent model constants,
// box, Verlet radii, integrator and thermostat); first / n / cells are the
// caller's.  Returns r_verlet in rv.  state_arrays: the desc must carry the
// Bounds of far pairs kept by their backbone sites (used where far_back is
// set): a backbone site may move by `back_skin` (the Verlet skin, or
// OXH_BACK_SKIN) from its place at the build before a rebuild; a far pair
// is kept while its sites are closer than the Debye-Hueckel cut-off plus
// twice that, plus a margin far above the FP32 error of wrapped sites in
// this box
void far_back_bounds(oxh_sys &S, double skin) {
	double back_skin = skin;
	if(const char *e = getenv("OXH_BACK_SKIN")) back_skin = atof(e);
	const double lmax = std::max(S.box[0], std::max(S.box[1], S.box[2]));
	const double margin = 1e-3 + std::ldexp(lmax, -18);
	const double r = S.dh_rc_d + 2.0 * back_skin + margin;
	S.rbf2 = (float) (r * r);
	S.sback2f = (float) (back_skin * back_skin);
	// sites of a close pair are within R_CLOSE + 2 skin + 2 backbone arms
	const double rg = std::max(r, (double) oxh::R_CLOSE + 2.0 * skin + 2.0 * 0.4814 + margin);
	S.rgate2f = (float) (rg * rg);
}

// state (pos, a1, a3) as well as the topology.  B (may be NULL) records
// the Bussi step alignment of a batch.
oxh_status sys_setup(oxh_batch *B, const oxh_system_desc &d, int s, oxh_sys &S, double &rv, bool state_arrays = true) {
	if(d.N <= 0 || !d.btype || !d.n3 || !d.n5 || (state_arrays && (!d.pos || !d.a1 || !d.a3)))
		return fail(OXH_ERR_ARG, "system %d: missing arrays or N <= 0", s);
	for(int k = 0; k < 3; k++)
		if(!(d.box[k] > 0)) return fail(OXH_ERR_ARG, "system %d: box side %d must be > 0", s, k);
	if(!(d.T > 0)) return fail(OXH_ERR_ARG, "system %d: T must be > 0", s);
	if(!(d.salt > 0)) return fail(OXH_ERR_ARG, "system %d: salt concentration must be > 0", s);
	if(d.interaction < OXH_INTERACTION_DNA2 || d.interaction > OXH_INTERACTION_RNA)
		return fail(OXH_ERR_ARG, "system %d: unknown interaction %d", s, d.interaction);
	if(d.interaction != OXH_INTERACTION_DNA2 && d.seq_stck)
		return fail(OXH_ERR_ARG, "system %d: the oxRNA models take the average-sequence parameters only", s);
	for(int i = 0; i < d.N; i++) {
		if(d.n3[i] >= d.N || d.n5[i] >= d.N || d.n3[i] < -1 || d.n5[i] < -1)
			return fail(OXH_ERR_ARG, "system %d: wrong topology for particle %d", s, i);
	}
	std::memset(&S, 0, sizeof(S));
	double rcut = 0;
	model_constants(d, S, rcut);
	for(int k = 0; k < 3; k++) {
		S.box[k] = d.box[k];
		S.inv_box[k] = 1.0 / d.box[k];
		S.boxf[k] = (float) d.box[k];
		S.inv_boxf[k] = (float) (1.0 / d.box[k]);
	}
	rv = rcut + 2.0 * d.verlet_skin;
	S.rv2 = rv * rv;
	S.skin2 = d.verlet_skin * d.verlet_skin;
	S.rclose2 = ((double) oxh::R_CLOSE + 2.0 * d.verlet_skin) * ((double) oxh::R_CLOSE + 2.0 * d.verlet_skin);
	{
		double lmax = std::max(d.box[0], std::max(d.box[1], d.box[2]));
		// |d r| <= 4 L 2^-24 per component after wrapping, rounding and the
		// nearest image; |d r^2| <= 2 |r| |d r| sqrt(3) + rounding of r^2
		double dr = 4.0 * lmax * std::ldexp(1.0, -24);
		S.eps2 = (float) (4.0 * (2.0 * rv * dr * std::sqrt(3.0) + 3.0 * dr * dr + 8.0 * rv * rv * std::ldexp(1.0, -24)));
		S.rv2f = (float) S.rv2;
		S.rclose2f = (float) S.rclose2;
	}
	S.small_box = 0;
	for(int k = 0; k < 3; k++)
		if(d.box[k] < 2.0 * (rv + 2.0 * 0.4814)) S.small_box = 1;
	far_back_bounds(S, d.verlet_skin);
	// integrator and thermostat (src/Backends/Thermostats/*.cpp)
	S.dt = d.dt;
	S.T = d.T;
	S.thermostat = d.thermostat;
	S.newtonian_steps = std::max(1, d.newtonian_steps);
	S.seed_lo = (uint32_t) (d.seed & 0xFFFFFFFFull);
	S.seed_hi = (uint32_t) (d.seed >> 32) ^ 0x6F78444Eu;
	S.pad_ = d.refresh_vel;
	if(d.thermostat == OXH_THERMOSTAT_LANGEVIN) {
		double D = 0, gamma = 0;
		if(d.diff_coeff > 0) {
			D = (double) (float) d.diff_coeff;
			gamma = d.T / D;
		}
		else if(d.gamma_trans > 0) {
			gamma = (double) (float) d.gamma_trans;
			D = d.T / gamma;
		}
		else return fail(OXH_ERR_ARG, "system %d: Langevin needs diff_coeff or gamma_trans", s);
		double Dr = 3. * D, gr = d.T / Dr;
		S.gamma_t = gamma;
		S.gamma_r = gr;
		S.resc_t = std::sqrt(2. * gamma * d.T / d.dt);
		S.resc_r = std::sqrt(2. * gr * d.T / d.dt);
	}
	else if(d.thermostat == OXH_THERMOSTAT_BROWNIAN) {
		double ns = S.newtonian_steps, dt = (double) (float) d.dt, pt = 0, D = 0;
		if(d.pt > 0) pt = (double) (float) d.pt;
		else if(d.diff_coeff > 0) {
			D = (double) (float) d.diff_coeff;
			pt = (2 * d.T * ns * dt) / (d.T * ns * dt + 2 * D);
		}
		else return fail(OXH_ERR_ARG, "system %d: brownian needs pt or diff_coeff", s);
		if(pt > 1.) return fail(OXH_ERR_ARG, "system %d: pt (%f) must be smaller than 1", s, pt);
		D = d.T * ns * dt * (1. / pt - 1. / 2.);
		S.pt = pt;
		S.pr = (2 * d.T * ns * dt) / (d.T * ns * dt + 2 * 3 * D);
		S.resc_brown = std::sqrt(d.T);
	}
	else if(d.thermostat == OXH_THERMOSTAT_BUSSI) {
		if(d.bussi_tau <= 0) return fail(OXH_ERR_ARG, "system %d: bussi_tau must be > 0", s);
		S.bussi_c = std::exp(-S.newtonian_steps / (double) d.bussi_tau);
		if(B) {
			B->any_bussi = true;
			if(B->bussi_ns == 0) B->bussi_ns = S.newtonian_steps;
			else if(B->bussi_ns != S.newtonian_steps) B->bussi_aligned = false;
		}
	}
	else if(d.thermostat != OXH_THERMOSTAT_NO) return fail(OXH_ERR_ARG, "system %d: unknown thermostat %d", s, d.thermostat);
	return OXH_OK;
}

// Cells of side >= r_verlet, at most ~2 particles per cell on average (the
// reference's cells_auto_optimisation, src/CUDA/Lists/
// CUDASimpleVerletList.cu:96-110); systems of at most 1024 particles (or
// fewer than 3 cells a side) use the all-pairs build.  max_cells > 0 caps
// the cell count (slots of a pool have a fixed cell budget: cells then
// get larger, which only adds candidates).  Updates the estimates of the
// largest cell occupancy and of the neighbours per particle.
void sys_cells(const oxh_system_desc &d, oxh_sys &S, double rv, int cell_first, int max_cells, int &max_occ,
		double &max_neigh_est) {
	double V = d.box[0] * d.box[1] * d.box[2];
	// half-size cells (edges >= r_verlet / 2, stencil +-2, at most 16 N
	// cells, at least 5 per side) or full-size ones (edges >= r_verlet,
	// stencil +-1, at most 2 N cells, at least 3 per side): half-size cells
	// where they hold 2 or more particles on average (fewer candidates per
	// particle), or where full-size ones do not fit
	int ncs[3][3];
	bool ok[3] = {false, false, false};
	long long ncell_of[3] = {0, 0, 0};
	for(int st = 1; st <= 2 && d.N > 1024; st++) {
		const double maxf = std::pow((st == 2 ? 16.0 : 2.0) * d.N / V, 1.0 / 3.0);
		int *nc = ncs[st];
		for(int k = 0; k < 3; k++) {
			nc[k] = (int) std::floor(d.box[k] * st / rv);
			nc[k] = std::min(nc[k], (int) std::ceil(maxf * d.box[k]));
		}
		if(max_cells > 0) {
			while((long long) nc[0] * nc[1] * nc[2] > max_cells) {
				int kmax = (nc[0] >= nc[1] && nc[0] >= nc[2]) ? 0 : (nc[1] >= nc[2] ? 1 : 2);
				nc[kmax]--;
			}
		}
		ok[st] = nc[0] >= 2 * st + 1 && nc[1] >= 2 * st + 1 && nc[2] >= 2 * st + 1;
		ncell_of[st] = (long long) nc[0] * nc[1] * nc[2];
	}
	int stencil = 0;
	if(ok[2] && (!ok[1] || d.N >= 2 * ncell_of[2])) stencil = 2;
	else if(ok[1]) stencil = 1;
	const bool cells_ok = stencil > 0;
	const int *nc = ncs[stencil];
	S.nc[0] = S.nc[1] = S.nc[2] = 0;
	S.cell_first = cell_first;
	S.ncell = 0;
	S.stencil = 0;
	if(cells_ok) {
		S.use_cells = 1;
		S.stencil = stencil;
		for(int k = 0; k < 3; k++) {
			S.nc[k] = nc[k];
			S.cell_w[k] = (float) (d.box[k] / nc[k]);
		}
		S.ncell = nc[0] * nc[1] * nc[2];
		// occupancy of the initial configuration
		std::vector<int> occ(S.ncell, 0);
		for(int i = 0; i < d.N; i++) {
			int c[3];
			for(int k = 0; k < 3; k++) {
				double sx = d.pos[3 * i + k] / d.box[k];
				int ck = (int) ((sx - std::floor(sx)) * nc[k]);
				c[k] = std::min(ck, nc[k] - 1);
			}
			int idx = (c[2] * nc[1] + c[1]) * nc[0] + c[0];
			max_occ = std::max(max_occ, ++occ[idx]);
		}
		// neighbours: the list sphere at a multiple of the fullest cell's
		// density (3x for full-size cells, 1.5x for the noisier half-size
		// ones); an underestimate only costs a regrow
		double cellvol = V / S.ncell;
		double est = 4.0 / 3.0 * M_PI * rv * rv * rv * ((stencil == 2 ? 1.5 : 3.0) * max_occ / cellvol);
		max_neigh_est = std::max(max_neigh_est, std::min(est, (double) d.N - 1));
	}
	else {
		S.use_cells = 0;
		max_neigh_est = std::max(max_neigh_est, (double) std::min(d.N - 1, 2048));
	}
}

// host staging of system s (desc d) that starts at batch index f; the
// arrays point at the system's first particle
void stage_system(const oxh_system_desc &d, int s, int f, double4 *hpos, double4 *hvel, double4 *hL, double4 *hq, int2 *hbonds,
		int2 *hinfo, int *horig) {
	for(int i = 0; i < d.N; i++) {
		hpos[i] = make_double4(d.pos[3 * i], d.pos[3 * i + 1], d.pos[3 * i + 2], 0.0);
		hvel[i] = d.vel ? make_double4(d.vel[3 * i], d.vel[3 * i + 1], d.vel[3 * i + 2], 0.0) : make_double4(0, 0, 0, 0);
		hL[i] = d.L ? make_double4(d.L[3 * i], d.L[3 * i + 1], d.L[3 * i + 2], 0.0) : make_double4(0, 0, 0, 0);
		double q[4];
		frame_to_quat(d.a1 + 3 * i, d.a3 + 3 * i, q);
		hq[i] = make_double4(q[0], q[1], q[2], q[3]);
		hbonds[i] = make_int2(d.n3[i] >= 0 ? f + d.n3[i] : -1, d.n5[i] >= 0 ? f + d.n5[i] : -1);
		hinfo[i] = make_int2(d.btype[i], s);
		horig[i] = i;
	}
}

// strand index of every particle: chains of n3 / n5 bonds, numbered in
// the order of their first particle (the reference's strand ids come from
// the topology file, TopologyParser.cpp:15-172; bonds define the same
// partition)
int strands_of(const oxh_system_desc &d, std::vector<int> &sid) {
	sid.assign(d.N, -1);
	int ns = 0;
	for(int i = 0; i < d.N; i++) {
		if(sid[i] >= 0) continue;
		// walk to the 5' end, then along n3
		int a = i, guard = 0;
		while(d.n5[a] >= 0 && d.n5[a] != i && guard++ < d.N) a = d.n5[a];
		for(int b = a, g = 0; b >= 0 && sid[b] < 0 && g <= d.N; b = d.n3[b], g++) sid[b] = ns;
		if(sid[i] < 0) sid[i] = ns;   // defensive: inconsistent bonds
		ns++;
	}
	return ns;
}

void initial_state(const oxh_system_desc &d, oxh_sys_state &st) {
	std::memset(&st, 0, sizeof(st));
	st.K_t = 0.5 * 3.0 * (d.N - 1) * d.T;
	st.K_r = 0.5 * 3.0 * d.N * d.T;
}

// reduction blocks: chunks of RED_CHUNK particles inside each system's
// current range; entries past the used ones are (-1, 0, 0)
void red_layout(const oxh_batch *B, std::vector<int3> &rb, std::vector<int2> &rsb, int nred_total) {
	rb.clear();
	rsb.assign(B->b.nsys, make_int2(0, 0));
	for(int s = 0; s < B->b.nsys; s++) {
		rsb[s].x = (int) rb.size();
		for(int a = B->first[s]; a < B->first[s] + B->n[s]; a += oxh::RED_CHUNK)
			rb.push_back(make_int3(s, a, std::min(a + oxh::RED_CHUNK, B->first[s] + B->n[s])));
		rsb[s].y = (int) rb.size() - rsb[s].x;
	}
	while((int) rb.size() < nred_total) rb.push_back(make_int3(-1, 0, 0));
}
}  // namespace

namespace {
template<typename T>
oxh_status download(oxh_batch *B, std::vector<T> &out, const T *src, int first, int n) {
	out.resize(n);
	HIPCHK(hipMemcpyAsync(out.data(), src + first, n * sizeof(T), hipMemcpyDeviceToHost, B->b.stream));
	return OXH_OK;
}
oxh_status check_sys(oxh_batch *B, int sys) {
	if(!B) return fail(OXH_ERR_ARG, "null batch");
	if(sys < 0 || sys >= B->b.nsys) return fail(OXH_ERR_ARG, "system %d out of range", sys);
	HIPCHK(hipSetDevice(B->b.device));
	return OXH_OK;
}
// moves particle i by -c * box (c: whole boxes per axis) and counts the
// move in its image counts, as OrthogonalBox::shift_particle does
// (src/Boxes/OrthogonalBox.cpp:67-74)
__global__ void k_shift(int N, const int4 *__restrict__ shift, const int2 *__restrict__ pinfo, const oxh_sys *__restrict__ systems,
		double4 *__restrict__ pos, double4 *__restrict__ rlist, int4 *__restrict__ img) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= N) return;
	const int4 c = shift[i];
	if(c.x == 0 && c.y == 0 && c.z == 0) return;
	const oxh_sys &S = systems[pinfo[i].y];
	const double dx = S.box[0] * (double) c.x, dy = S.box[1] * (double) c.y, dz = S.box[2] * (double) c.z;
	double4 p = pos[i], r = rlist[i];
	p.x -= dx;
	p.y -= dy;
	p.z -= dz;
	r.x -= dx;
	r.y -= dy;
	r.z -= dz;
	pos[i] = p;
	rlist[i] = r;   // relative positions are unchanged: lists stay valid
	int4 m = img[i];
	m.x += c.x;
	m.y += c.y;
	m.z += c.z;
	img[i] = m;
}

// Moves every strand whose centre of mass left the box back by whole box
// lengths (SimBackend::fix_diffusion, src/Backends/SimBackend.cpp:788-882,
// with OrthogonalBox::shift_particle, src/Boxes/OrthogonalBox.cpp:67-74):
// centres of mass in FP64 in the original particle order, the same shift
// for every particle of a strand; positions and list positions move
// together.  Orientations are unit quaternions renormalised every step, so
// the reference's re-orthonormalisation has nothing to do.  If the
// potential energy changes by more than 1e-6 of itself (as the reference
// checks) the shift is undone.
oxh_status fix_diffusion_impl(oxh_batch *B, int *n_moved) {
	oxh::Batch &b = B->b;
	std::vector<double4> p;
	std::vector<int> o;
	TRY(download(B, p, b.pos, 0, b.N));
	TRY(download(B, o, b.orig, 0, b.N));
	std::vector<double> E0(4 * b.nsys);
	TRY(sync_and_check(B));
	std::vector<int4> shift(b.N, make_int4(0, 0, 0, 0));
	int moved = 0;
	for(int s = 0; s < b.nsys; s++) {
		const int f = B->first[s], n = B->n[s], ns = B->n_strands[s];
		if(n == 0 || ns == 0) continue;
		std::vector<int> where(n);
		for(int k = 0; k < n; k++) where[o[f + k]] = f + k;
		std::vector<double> com(3 * ns, 0.0);
		std::vector<int> cnt(ns, 0);
		const std::vector<int> &sid = B->strand_of[s];
		for(int i = 0; i < n; i++) {
			const double4 x = p[where[i]];
			com[3 * sid[i]] += x.x;
			com[3 * sid[i] + 1] += x.y;
			com[3 * sid[i] + 2] += x.z;
			cnt[sid[i]]++;
		}
		const oxh_sys &S = b.h_sys[s];
		std::vector<int> d(3 * ns, 0);
		for(int m = 0; m < ns; m++) {
			bool any = false;
			for(int c = 0; c < 3; c++) {
				const double cm = com[3 * m + c] / cnt[m];
				d[3 * m + c] = (int) std::floor(cm / S.box[c]);
				any |= d[3 * m + c] != 0;
			}
			moved += any;
		}
		for(int i = 0; i < n; i++) {
			const int m = sid[i];
			shift[where[i]] = make_int4(d[3 * m], d[3 * m + 1], d[3 * m + 2], 0);
		}
	}
	if(n_moved) *n_moved = moved;
	if(moved == 0) return OXH_OK;
	int4 *dshift = nullptr;
	TRY(dalloc(B, &dshift, b.N));
	auto apply = [&](double sign) -> oxh_status {
		if(sign < 0)
			for(auto &x : shift) x = make_int4(-x.x, -x.y, -x.z, 0);
		HIPCHK(hipMemcpyAsync(dshift, shift.data(), b.N * sizeof(int4), hipMemcpyHostToDevice, b.stream));
		k_shift<<<(b.N + 255) / 256, 256, 0, b.stream>>>(b.N, dshift, b.pinfo, b.d_sys, b.pos, b.rlist, b.img);
		oxh::launch_prepare_mirror(b.stream, b);
		enqueue_forces(B);
		return sync_and_check(B);
	};
	std::vector<double> before(4 * b.nsys), after(4 * b.nsys);
	oxh_status st = oxh_batch_energies(B, before.data());
	if(st == OXH_OK) st = apply(1.0);
	if(st == OXH_OK) st = oxh_batch_energies(B, after.data());
	bool undo = false;
	for(int s = 0; st == OXH_OK && s < b.nsys; s++)
		undo |= std::fabs(after[4 * s] - before[4 * s]) > 1e-6 * std::fabs(before[4 * s]);
	if(st == OXH_OK && undo) {
		st = apply(-1.0);
		if(n_moved) *n_moved = 0;
	}
	(void) hipFree(dshift);
	B->device_bytes -= (size_t) b.N * sizeof(int4);
	return st;
}

// Picks the force kernels' launch shapes by timing every candidate on this
// batch's own lists: threads per nucleotide and block size of the close
// (+ bonded) kernel, and whether the far pairs run inline or in their own
// low-register kernel (and its shape).  The best shapes depend on N (waves
// to fill 256 CUs) and on the close / far neighbour counts.
oxh_status autotune_forces(oxh_batch *B) {
	oxh::Batch &b = B->b;
	const char *edge_env = getenv("OXH_EDGE");
	const char *pack_env = getenv("OXH_PACK");
	if(const char *env = getenv("OXH_FORCE_LANES")) {
		b.lanes = atoi(env);
		b.pack = !b.fp64 && b.lanes > 1 && pack_env && atoi(pack_env) == 1;
		if(const char *f = getenv("OXH_FAR_LANES")) b.far_lanes = atoi(f);
		b.edge = b.ecap > 0 && edge_env && atoi(edge_env) == 1;
		if(const char *g = getenv("OXH_EDGE_GROUP")) b.edge_group = atoi(g);
		if(b.ecap > 0 && !b.edge) TRY(alloc_edges(B, 0));
		return OXH_OK;
	}
	// the workgroup engine fixes its own lanes per system
	if(B->persist) return OXH_OK;
	hipEvent_t e0, e1;
	HIPCHK(hipEventCreate(&e0));
	HIPCHK(hipEventCreate(&e1));
	auto timeit = [&](float &ms) -> oxh_status {
		oxh::launch_forces(b.stream, b, false);
		HIPCHK(hipEventRecord(e0, b.stream));
		for(int k = 0; k < 3; k++) oxh::launch_forces(b.stream, b, false);
		HIPCHK(hipEventRecord(e1, b.stream));
		HIPCHK(hipEventSynchronize(e1));
		HIPCHK(hipEventElapsedTime(&ms, e0, e1));
		return OXH_OK;
	};
	const int lanes[] = {1, 2, 4, 8, 16, 32};
	const int blocks[] = {64, 128, 256};
	float best = 1e30f;
	if(b.model) {
		// oxRNA: one FP64 kernel over the whole row, blocks of 64
		b.far_lanes = 0;
		b.pack = false;
		int bl = 4;
		for(int L : {1, 2, 4, 8, 16}) {
			b.lanes = L;
			float ms;
			TRY(timeit(ms));
			if(ms < best) {
				best = ms;
				bl = L;
			}
		}
		(void) hipEventDestroy(e0);
		(void) hipEventDestroy(e1);
		b.lanes = bl;
		enqueue_forces(B);
		return sync_and_check(B);
	}
	if(b.fp64) {
		// all-FP64 mode: close kernel in blocks of 64 with far pairs inline,
		// then far pairs in their own kernel (blocks of 256)
		b.force_block = 64;
		b.far_lanes = 0;
		int bl64 = 8, bf64 = 0;
		for(int L : lanes) {
			if(L == 1) continue;
			b.lanes = L;
			float ms;
			TRY(timeit(ms));
			if(ms < best) {
				best = ms;
				bl64 = L;
			}
		}
		b.lanes = bl64;
		for(int FL : {4, 8, 16, 32}) {
			for(int L : {2, 4, 8}) {
				b.lanes = L;
				b.far_lanes = FL;
				float ms;
				TRY(timeit(ms));
				if(ms < best) {
					best = ms;
					bl64 = L;
					bf64 = FL;
				}
			}
		}
		(void) hipEventDestroy(e0);
		(void) hipEventDestroy(e1);
		b.lanes = bl64;
		b.far_lanes = bf64;
		enqueue_forces(B);
		return sync_and_check(B);
	}
	int bl = b.lanes, bb = b.force_block, bfl = 0, bfb = 256;
	bool bp = false;
	// close kernel shapes: lanes x workgroup size, and lanes with the
	// packed form (OXH_PACK=0/1 keeps one form), at the current far shape
	auto scan_close = [&]() -> oxh_status {
		for(int pk = 0; pk < 2; pk++) {
			if(pack_env && atoi(pack_env) != pk) continue;
			b.pack = pk == 1;
			for(int L : lanes)
				for(int bs : blocks) {
					if(b.pack && (bs != oxh::PACK_BS || L == 1)) continue;
					b.lanes = L;
					b.force_block = bs;
					float ms;
					TRY(timeit(ms));
					if(ms < best) {
						best = ms;
						bl = L;
						bb = bs;
						bp = b.pack;
					}
				}
		}
		b.lanes = bl;
		b.force_block = bb;
		b.pack = bp;
		return OXH_OK;
	};
	// 1) close kernel shape with the far pairs inline
	b.far_lanes = 0;
	TRY(scan_close());
	// 2) far pairs in their own kernel, close kernel at its best shape; if
	// that wins, the close shape again next to the far kernel
	for(int L : lanes)
		for(int bs : blocks) {
			b.far_lanes = L;
			b.far_block = bs;
			float ms;
			TRY(timeit(ms));
			if(ms < best) {
				best = ms;
				bfl = L;
				bfb = bs;
			}
		}
	b.far_lanes = bfl;
	b.far_block = bfb;
	if(bfl != 0) TRY(scan_close());
	// 3) the half-list form of the close pairs: particles per wave of the
	// pair kernel x lanes per particle of the gather (far pairs inline, or
	// in the far kernel picked above)
	if(b.ecap > 0) {
		const bool forced = edge_env && atoi(edge_env) == 1;
		const int groups[] = {4, 8, 16, 32, 64};
		float best_edge = 1e30f;
		int el = 8, eg = 16, efl = bfl;
		b.edge = true;
		for(int fl : {0, bfl}) {
			b.far_lanes = fl;
			for(int G : groups)
				for(int L : lanes) {
					b.edge_group = G;
					b.lanes = L;
					float ms;
					TRY(timeit(ms));
					if(ms < best_edge) {
						best_edge = ms;
						el = L;
						eg = G;
						efl = fl;
					}
				}
			if(bfl == 0) break;
		}
		if(forced || best_edge < best) {
			b.lanes = el;
			b.edge_group = eg;
			b.far_lanes = efl;
		}
		else {
			b.edge = false;
			b.lanes = bl;
			b.force_block = bb;
			b.far_lanes = bfl;
			TRY(alloc_edges(B, 0));
		}
	}
	(void) hipEventDestroy(e0);
	(void) hipEventDestroy(e1);
	// forces of the current state with the chosen shapes
	enqueue_forces(B);
	return sync_and_check(B);
}

}  // namespace

extern "C" {

void oxh_system_desc_defaults(oxh_system_desc *d) {
	std::memset(d, 0, sizeof(*d));
	d->T = 0.1;
	d->salt = 0.5;
	d->dh_half_charged_ends = 1;
	d->dh_lambda = -1;
	d->dh_strength = -1;
	d->dh_rhigh = -1;
	d->hb_multiplier = 1.0;
	d->max_backbone_force = -1;
	d->max_backbone_force_far = -1;
	d->verlet_skin = 0.05;
	d->dt = 0.003;
	d->thermostat = OXH_THERMOSTAT_NO;
	d->diff_coeff = -1;
	d->gamma_trans = -1;
	d->pt = -1;
	d->newtonian_steps = 1;
	d->bussi_tau = 1;
	d->seed = 1;
}

const char *oxh_last_error(void) { return g_err.c_str(); }

int oxh_version(void) { return 10000; }

oxh_status oxh_device_count(int *n) {
	HIPCHK(hipGetDeviceCount(n));
	return OXH_OK;
}

oxh_status oxh_batch_destroy(oxh_batch *B) {
	if(!B) return OXH_OK;
	oxh::Batch &b = B->b;
	(void) hipSetDevice(b.device);
	if(b.stream) (void) hipStreamSynchronize(b.stream);
	if(b.graph_exec) (void) hipGraphExecDestroy(b.graph_exec);
	void *ptrs[] = {b.pos, b.vel, b.Lm, b.q64, b.rlist, b.q32, b.bpos, b.F, b.T, b.bonds, b.pinfo, b.orig, b.pstep, b.nbr, b.nnbr,
			b.cell_count, b.cell_items, b.cell_owner, b.pos32, b.d_sys, b.d_state, b.red_blocks, b.red_sys_blocks, b.red_partial, b.red_out, b.red_done,
			b.esplit, b.stress, b.hbcount, b.keys, b.keys_alt, b.perm, b.perm_alt, b.inv_perm, b.radix_hist, b.scratch, B->bk_pos, B->bk_vel,
			B->bk_Lm, B->bk_q64, B->bk_F, B->bk_T, B->bk_pstep, B->bk_state, B->bk_rlist, b.limit, b.any_rebuild, b.F64, b.T64, B->bk_F64, B->bk_T64, b.eown, b.ein, b.ecount, b.E1, b.E2, b.E3, b.B1, b.B2, b.B3, b.cell_raw, b.cell_of, b.back64, b.img, b.slot_of, b.ext_off, b.ext_rec,
			b.ext_grp, b.ext_grp_sys, b.ext_idx, b.ext_com, b.blist, B->bk_blist};
	for(void *p : ptrs)
		if(p) (void) hipFree(p);
	if(b.stream) (void) hipStreamDestroy(b.stream);
	delete B;
	return OXH_OK;
}

}  // extern "C"

namespace {

// A batch of n_sys systems (descs), or a slot pool (capacity != nullptr):
// slot s reserves capacity[s] particles and cell_budget[s] cells and holds
// descs[s] if given and non-empty, else nothing.
oxh_status create_impl(int device, int n_sys, const oxh_system_desc *descs, const int *capacity, const int *cell_budget,
		oxh_batch **out) {
	*out = nullptr;
	int ndev = 0;
	HIPCHK(hipGetDeviceCount(&ndev));
	if(device < 0 || device >= ndev) return fail(OXH_ERR_ARG, "device %d not available (%d devices)", device, ndev);
	HIPCHK(hipSetDevice(device));

	oxh_batch *B = new oxh_batch();
	oxh::Batch &b = B->b;
	b.device = device;
	b.nsys = n_sys;
	B->pool = capacity != nullptr;
	auto bail = [&](oxh_status s) {
		std::string msg = g_err;
		oxh_batch_destroy(B);
		g_err = msg;
		return s;
	};
	if(descs) {
		b.fp64 = descs[0].backend_precision == 1;
		for(int s = 0; s < n_sys; s++) {
			if(descs[s].backend_precision != 0 && descs[s].backend_precision != 1)
				return bail(fail(OXH_ERR_ARG, "system %d: backend_precision must be 0 (mixed) or 1 (double)", s));
			if((descs[s].backend_precision == 1) != b.fp64)
				return bail(fail(OXH_ERR_ARG, "all systems of a batch need the same backend_precision"));
			if(descs[s].interaction != descs[0].interaction)
				return bail(fail(OXH_ERR_ARG, "all systems of a batch need the same interaction"));
		}
		b.model = descs[0].interaction == OXH_INTERACTION_DNA2 ? 0 : 1;
		if(b.model && B->pool) return bail(fail(OXH_ERR_ARG, "slot pools run the oxDNA2 model only"));
	}
#define TRYB(x)                               \
	do {                                      \
		oxh_status s_ = (x);                  \
		if(s_ != OXH_OK) return bail(s_);     \
	} while(0)
	if(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking) != hipSuccess) return bail(fail(OXH_ERR_HIP, "stream creation failed"));

	// ---- per system parameters
	b.h_sys.resize(n_sys);
	B->first.resize(n_sys);
	B->n.resize(n_sys);
	int N = 0, ncell = 0, max_occ = 1;
	double max_neigh_est = 16;
	b.min_newtonian = 1 << 30;
	if(B->pool) {
		B->cap.assign(capacity, capacity + n_sys);
		B->cell_budget.assign(cell_budget, cell_budget + n_sys);
		B->target.assign(n_sys, 0);
		B->done_steps.assign(n_sys, 0);
		B->h_limit.assign(n_sys, 0);
	}
	for(int s = 0; s < n_sys; s++) {
		oxh_sys &S = b.h_sys[s];
		const bool has = descs && descs[s].N > 0;
		if(has) {
			double rv = 0;
			TRYB(sys_setup(B, descs[s], s, S, rv));
			sys_cells(descs[s], S, rv, ncell, B->pool ? cell_budget[s] : 0, max_occ, max_neigh_est);
		}
		else {
			std::memset(&S, 0, sizeof(S));
			S.cell_first = ncell;
		}
		S.first = N;
		S.n = has ? descs[s].N : 0;
		B->first[s] = N;
		B->n[s] = S.n;
		ncell += B->pool ? cell_budget[s] : S.ncell;
		N += B->pool ? capacity[s] : S.n;
	}
	if(B->pool) B->bussi_aligned = false;   // slots start at different times
	// the oxRNA path runs on systems without cell lists: a 24,576-nt oxRNA2
	// system with cells left the GPU faulted, with no cause found in this
	// path's code; until one is, such systems are refused
	for(int s = 0; s < n_sys && b.model; s++)
		if(b.h_sys[s].use_cells)
			return bail(fail(OXH_ERR_ARG, "system %d: the oxRNA models take systems without cell lists (at most 1024 nucleotides)", s));
	b.N = N;
	b.nstride = N;
	b.ncell_total = ncell;
	b.any_cells = ncell > 0;
	// cell-centric list build where cells hold several particles each
	// (every system with cells, not a pool; the rows are the same either
	// way; OXH_CELL_BUILD=0 / 1 forbids / forces it)
	{
		bool all_cells = !B->pool && ncell > 0;
		for(int s = 0; s < n_sys; s++) all_cells &= b.h_sys[s].use_cells != 0;
		const char *e = getenv("OXH_CELL_BUILD");
		b.cell_build = all_cells && (e ? atoi(e) == 1 : N >= 4 * ncell);
	}
	b.cell_cap = std::max(8, 3 * max_occ);
	b.max_neigh = std::max(8, (int) std::ceil(max_neigh_est));
	if(b.max_neigh > N) b.max_neigh = std::max(1, N);

	// ---- device buffers
	TRYB(dalloc(B, &b.pos, N));
	TRYB(dalloc(B, &b.vel, N));
	TRYB(dalloc(B, &b.Lm, N));
	TRYB(dalloc(B, &b.q64, N));
	TRYB(dalloc(B, &b.rlist, N));
	TRYB(dalloc(B, &b.q32, N));
	TRYB(dalloc(B, &b.bpos, N));
	TRYB(dalloc(B, &b.F, N));
	TRYB(dalloc(B, &b.T, N));
	TRYB(dalloc(B, &b.bonds, N));
	TRYB(dalloc(B, &b.pinfo, N));
	TRYB(dalloc(B, &b.orig, N));
	TRYB(dalloc(B, &b.slot_of, N));
	TRYB(dalloc(B, &b.img, N));
	TRYB(dalloc(B, &b.pstep, N));
	TRYB(dalloc(B, &b.nnbr, N));
	TRYB(dalloc(B, &b.nbr, (size_t) N * b.max_neigh));
	TRYB(dalloc(B, &b.cell_count, std::max(ncell, 1)));
	TRYB(dalloc(B, &b.cell_owner, std::max(ncell, 1)));
	TRYB(dalloc(B, &b.pos32, N));
	TRYB(alloc_cell_slots(B, b.cell_cap));
	TRYB(dalloc(B, &b.cell_of, N));
	TRYB(dalloc(B, &b.d_sys, n_sys));
	TRYB(dalloc(B, &b.d_state, n_sys));
	TRYB(dalloc(B, &b.esplit, (size_t) 8 * N));
	TRYB(dalloc(B, &b.stress, (size_t) 6 * N));
	TRYB(dalloc(B, &b.hbcount, N));
	TRYB(dalloc(B, &b.keys, N));
	TRYB(dalloc(B, &b.keys_alt, N));
	TRYB(dalloc(B, &b.perm, N));
	TRYB(dalloc(B, &b.perm_alt, N));
	TRYB(dalloc(B, &b.inv_perm, N));
	TRYB(dalloc(B, &b.radix_hist, (size_t) 256 * ((N + 4095) / 4096)));
	b.scratch_bytes = (size_t) std::max(N, 2 * n_sys) * sizeof(double4);
	TRYB(dalloc(B, (char **) &b.scratch, b.scratch_bytes));
	TRYB(dalloc(B, &B->bk_pos, N));
	TRYB(dalloc(B, &B->bk_vel, N));
	TRYB(dalloc(B, &B->bk_Lm, N));
	TRYB(dalloc(B, &B->bk_q64, N));
	TRYB(dalloc(B, &B->bk_F, N));
	TRYB(dalloc(B, &B->bk_T, N));
	if(b.fp64) {
		TRYB(dalloc(B, &b.F64, N));
		TRYB(dalloc(B, &b.T64, N));
		TRYB(dalloc(B, &b.back64, N));
		TRYB(dalloc(B, &B->bk_F64, N));
		TRYB(dalloc(B, &B->bk_T64, N));
	}
	TRYB(dalloc(B, &B->bk_pstep, N));
	TRYB(dalloc(B, &B->bk_state, n_sys));

	// reduction blocks: chunks of RED_CHUNK particles inside one system
	std::vector<int3> rb;
	std::vector<int2> rsb;
	int nred_total = 0;   // pools: enough blocks for every slot at capacity
	if(B->pool)
		for(int s = 0; s < n_sys; s++) nred_total += (capacity[s] + oxh::RED_CHUNK - 1) / oxh::RED_CHUNK;
	red_layout(B, rb, rsb, nred_total);
	b.nred = std::max(1, (int) rb.size());
	if(rb.empty()) rb.push_back(make_int3(-1, 0, 0));
	TRYB(dalloc(B, &b.red_blocks, b.nred));
	TRYB(dalloc(B, &b.red_sys_blocks, n_sys));
	TRYB(dalloc(B, &b.red_partial, (size_t) b.nred * oxh::RED_W));
	TRYB(dalloc(B, &b.red_out, (size_t) n_sys * oxh::RED_W));
	TRYB(dalloc(B, &b.red_done, n_sys));

	// ---- host staging
	std::vector<double4> hpos(N), hvel(N), hL(N), hq(N);
	std::vector<int2> hbonds(N, make_int2(-1, -1)), hinfo(N);
	std::vector<int> horig(N), howner(std::max(ncell, 1));
	std::vector<oxh_sys_state> hst(n_sys);
	B->strand_of.assign(n_sys, std::vector<int>());
	B->n_strands.assign(n_sys, 0);
	for(int s = 0; s < n_sys; s++) {
		static const oxh_system_desc empty = {};
		const oxh_system_desc &d = descs ? descs[s] : empty;
		const oxh_sys &S = b.h_sys[s];
		const int f = B->first[s];
		int nres = S.ncell;   // cells owned by this system
		if(B->pool) nres = B->cell_budget[s];
		for(int c = 0; c < nres; c++) howner[S.cell_first + c] = s;
		for(int i = f; i < f + (B->pool ? B->cap[s] : B->n[s]); i++) hinfo[i] = make_int2(0, s);   // padding belongs to its slot
		if(!descs || d.N <= 0) continue;
		stage_system(d, s, f, hpos.data() + f, hvel.data() + f, hL.data() + f, hq.data() + f, hbonds.data() + f, hinfo.data() + f,
				horig.data() + f);
		initial_state(d, hst[s]);
		B->n_strands[s] = strands_of(d, B->strand_of[s]);
	}
	hipStream_t st = b.stream;
	auto up = [&](void *dst, const void *src, size_t bytes) {
		return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st) == hipSuccess;
	};
	bool ok = up(b.pos, hpos.data(), N * sizeof(double4)) && up(b.vel, hvel.data(), N * sizeof(double4)) &&
			up(b.Lm, hL.data(), N * sizeof(double4)) && up(b.q64, hq.data(), N * sizeof(double4)) &&
			up(b.bonds, hbonds.data(), N * sizeof(int2)) && up(b.pinfo, hinfo.data(), N * sizeof(int2)) &&
			up(b.orig, horig.data(), N * sizeof(int)) && up(b.cell_owner, howner.data(), howner.size() * sizeof(int)) &&
			up(b.d_sys, b.h_sys.data(), n_sys * sizeof(oxh_sys)) && up(b.d_state, hst.data(), n_sys * sizeof(oxh_sys_state)) &&
			up(b.red_blocks, rb.data(), rb.size() * sizeof(int3)) && up(b.red_sys_blocks, rsb.data(), n_sys * sizeof(int2));
	if(!ok) return bail(fail(OXH_ERR_HIP, "upload failed"));
	if(hipMemsetAsync(b.pstep, 0, N * sizeof(long long), st) != hipSuccess) return bail(fail(OXH_ERR_HIP, "memset failed"));
	if(hipMemsetAsync(b.img, 0, N * sizeof(int4), st) != hipSuccess) return bail(fail(OXH_ERR_HIP, "memset failed"));
	if(hipMemsetAsync(b.red_done, 0, n_sys * sizeof(int), st) != hipSuccess) return bail(fail(OXH_ERR_HIP, "memset failed"));
	if(hipMemsetAsync(b.cell_count, 0, std::max(ncell, 1) * sizeof(int), st) != hipSuccess) return bail(fail(OXH_ERR_HIP, "memset failed"));
	if(hipMemsetAsync(b.F, 0, N * sizeof(float4), st) != hipSuccess) return bail(fail(OXH_ERR_HIP, "memset failed"));
	if(hipMemsetAsync(b.T, 0, N * sizeof(float4), st) != hipSuccess) return bail(fail(OXH_ERR_HIP, "memset failed"));
	if(b.fp64 && (hipMemsetAsync(b.F64, 0, N * sizeof(double4), st) != hipSuccess ||
			hipMemsetAsync(b.T64, 0, N * sizeof(double4), st) != hipSuccess))
		return bail(fail(OXH_ERR_HIP, "memset failed"));

	// refresh_vel systems: draw on the device, keep the others as uploaded
	bool any_refresh = false;
	for(int s = 0; s < n_sys && descs; s++) any_refresh |= descs[s].N > 0 && descs[s].refresh_vel != 0;
	if(any_refresh) {
		std::vector<double4> keep_v = hvel, keep_L = hL;
		oxh::launch_refresh_vel(st, b, 0, N);
		std::vector<double4> nv(N), nL(N);
		(void) hipMemcpyAsync(nv.data(), b.vel, N * sizeof(double4), hipMemcpyDeviceToHost, st);
		(void) hipMemcpyAsync(nL.data(), b.Lm, N * sizeof(double4), hipMemcpyDeviceToHost, st);
		if(hipStreamSynchronize(st) != hipSuccess) return bail(fail(OXH_ERR_HIP, "refresh_vel failed"));
		for(int s = 0; s < n_sys; s++) {
			if(descs[s].refresh_vel) continue;
			for(int i = B->first[s]; i < B->first[s] + B->n[s]; i++) {
				nv[i] = keep_v[i];
				nL[i] = keep_L[i];
			}
		}
		(void) up(b.vel, nv.data(), N * sizeof(double4));
		(void) up(b.Lm, nL.data(), N * sizeof(double4));
	}
	oxh::launch_prepare_mirror(st, b);
	oxh::launch_slot_of(st, b);
	// the workgroup engine takes systems of at most one particle per thread
	if(const char *x = getenv("OXH_XCD")) b.xcd = atoi(x) != 0;
	if(const char *p = getenv("OXH_PERSIST_BLOCK")) B->persist_block = atoi(p) == 256 ? 256 : 512;
	B->persist = true;
	for(int s = 0; s < n_sys; s++) {
		const oxh_sys &S = b.h_sys[s];
		const int reach = B->pool ? B->cap[s] : S.n;
		B->persist &= !S.use_cells && reach <= std::min(oxh::PERSIST_MAX_N, B->persist_block);
	}
	if(const char *p = getenv("OXH_PERSIST")) B->persist &= atoi(p) != 0;
	B->persist &= !b.fp64;   // the workgroup engine computes in the mixed form
	B->persist &= !b.model;  // and the oxDNA2 terms
	// systems (slots) small enough are held in LDS by the workgroup engine
	if(B->persist) {
		int nmax = 0;
		for(int s = 0; s < n_sys; s++) nmax = std::max(nmax, B->pool ? B->cap[s] : b.h_sys[s].n);
		if(nmax <= oxh::PERSIST_LDS_MAX_N) b.persist_lds_n = nmax;
		if(const char *p = getenv("OXH_PERSIST_LDS"))
			if(atoi(p) == 0) b.persist_lds_n = 0;
	}
	// half-list form of the step's close pairs: grid-engine batches in the
	// mixed form, timed against the full list at upload (OXH_EDGE=0 / 1
	// forbids / forces it)
	{
		const char *e = getenv("OXH_EDGE");
		if(!B->pool && !B->persist && !b.fp64 && !b.model && !(e && atoi(e) == 0)) TRYB(alloc_edges(B, std::min(b.max_neigh, 64)));
	}
	// far pairs kept by their backbone sites: grid-engine batches and
	// pools, systems with cells in boxes wide enough that backbone sites
	// take the centres' nearest image (a pool's slots decide at admission;
	// OXH_FAR_BACK=0 keeps every far pair within r_verlet)
	{
		const char *e = getenv("OXH_FAR_BACK");
		const bool allowed = !B->persist && !b.model && !(e && atoi(e) == 0);
		bool any = false;
		for(int s = 0; s < n_sys; s++) {
			oxh_sys &S = b.h_sys[s];
			S.far_back = (allowed && !B->pool && S.use_cells && !S.small_box) ? 1 : 0;
			any |= S.far_back != 0;
		}
		if(any || (allowed && B->pool && ncell > 0)) {
			TRYB(dalloc(B, &b.blist, N));
			TRYB(dalloc(B, &B->bk_blist, N));
		}
		if(any && !up(b.d_sys, b.h_sys.data(), n_sys * sizeof(oxh_sys))) return bail(fail(OXH_ERR_HIP, "system upload failed"));
	}
	TRYB(dalloc(B, &B->bk_rlist, N));
	TRYB(dalloc(B, &b.any_rebuild, 1));
	if(hipMemsetAsync(b.any_rebuild, 0, sizeof(int), st) != hipSuccess) return bail(fail(OXH_ERR_HIP, "memset failed"));
	if(B->pool) TRYB(dalloc(B, &b.limit, n_sys));
	oxh_status s = build_all(B);
	if(s != OXH_OK) return bail(s);
	if(!B->pool) {
		s = autotune_forces(B);
		if(s != OXH_OK) return bail(s);
	}
	else if(const char *env = getenv("OXH_FORCE_LANES")) {
		b.lanes = atoi(env);
		b.far_lanes = 0;
		if(const char *f = getenv("OXH_FAR_LANES")) b.far_lanes = atoi(f);
	}
	// graph chunk: enough steps to amortise the launch, at most ~2k kernels
	B->chunk = 64;
	if(const char *g = getenv("OXH_GRAPH")) B->use_graph = atoi(g) != 0;
	if(B->any_bussi && B->bussi_aligned) B->chunk = std::max(B->bussi_ns, (64 / B->bussi_ns) * B->bussi_ns);
	*out = B;
	return OXH_OK;
#undef TRYB
}

}  // namespace

extern "C" {

oxh_status oxh_batch_create(int device, int n_sys, const oxh_system_desc *descs, oxh_batch **out) {
	if(!out || n_sys <= 0 || !descs) return fail(OXH_ERR_ARG, "oxh_batch_create: bad arguments");
	for(int s = 0; s < n_sys; s++)
		if(descs[s].N <= 0) return fail(OXH_ERR_ARG, "system %d: N must be > 0", s);
	return create_impl(device, n_sys, descs, nullptr, nullptr, out);
}

oxh_status oxh_batch_update_lists(oxh_batch *B, int force) {
	if(!B) return fail(OXH_ERR_ARG, "null batch");
	HIPCHK(hipSetDevice(B->b.device));
	for(int attempt = 0; attempt < 12; attempt++) {
		enqueue_lists(B, force != 0);
		k_clear_need<<<(B->b.nsys + 63) / 64, 64, 0, B->b.stream>>>(B->b.nsys, B->b.d_state, 0, B->b.any_rebuild);
		std::vector<oxh_sys_state> st;
		TRY(read_states(B, st));
		bool overflow = false;
		for(auto &x : st) overflow |= (x.error & oxh::ERR_OVERFLOW) != 0;
		if(!overflow) return OXH_OK;
		TRY(grow_lists(B, st));
		force = 1;
	}
	return fail(OXH_ERR_SYSTEM, "neighbour list keeps overflowing");
}

oxh_status oxh_batch_compute_forces(oxh_batch *B) {
	if(!B) return fail(OXH_ERR_ARG, "null batch");
	HIPCHK(hipSetDevice(B->b.device));
	enqueue_forces(B);
	return sync_and_check(B);
}

}  // extern "C"

namespace {

// The caller's state (upload order: particle k of system s at index
// first_s + k) into the batch's slots, with the FP32 mirror and the skin
// check against the positions of each system's last list build.
__global__ void k_load_caller_state(int N, const double4 *__restrict__ upos, const double4 *__restrict__ uq, const int *__restrict__ orig,
		const int2 *__restrict__ pinfo, const int2 *__restrict__ bonds, const oxh_sys *__restrict__ systems,
		const double4 *__restrict__ rlist, double4 *__restrict__ pos, double4 *__restrict__ q64, float4 *__restrict__ q32,
		float4 *__restrict__ bpos, double4 *__restrict__ back64, oxh_sys_state *__restrict__ st, int *__restrict__ any,
		const float4 *__restrict__ blist) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= N) return;
	const int s = pinfo[i].y;
	const oxh_sys &S = systems[s];
	if(oxh::padding(i, S)) return;
	const int u = S.first + orig[i];
	double4 r = upos[u];
	r.w = 0.0;
	const double4 q = uq[u];
	pos[i] = r;
	q64[i] = q;
	const float4 bp = oxh::write_mirror(i, r, q, S, bonds[i], q32, bpos, back64);
	const double4 rl = rlist[i];
	const double dx = r.x - rl.x, dy = r.y - rl.y, dz = r.z - rl.z;
	if(dx * dx + dy * dy + dz * dz > S.skin2 || (S.far_back && blist && oxh::back_moved(bp, blist[i], S))) {
		st[s].need_rebuild = 1;
		*any = 1;
	}
}

// caller's F / T (upload order) += the batch's forces and torques, .w
// included (energy share, HB energy share)
template<class FT>
__global__ void k_add_to_caller(int N, const FT *__restrict__ F, const FT *__restrict__ T, const int *__restrict__ orig,
		const int2 *__restrict__ pinfo, const oxh_sys *__restrict__ systems, FT *__restrict__ uF, FT *__restrict__ uT) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= N) return;
	const oxh_sys &S = systems[pinfo[i].y];
	if(oxh::padding(i, S)) return;
	const int u = S.first + orig[i];
	FT a = uF[u];
	const FT f = F[i];
	a.x += f.x;
	a.y += f.y;
	a.z += f.z;
	a.w += f.w;
	uF[u] = a;
	FT b = uT[u];
	const FT t = T[i];
	b.x += t.x;
	b.y += t.y;
	b.z += t.z;
	b.w += t.w;
	uT[u] = b;
}

}  // namespace

extern "C" {

oxh_status oxh_interaction_compute_forces(oxh_batch *B, const double *d_pos, const double *d_quat, void *d_forces, void *d_torques,
		void *stream) {
	if(!B || !d_pos || !d_quat || !d_forces || !d_torques) return fail(OXH_ERR_ARG, "oxh_interaction_compute_forces: bad arguments");
	if(B->pool) return fail(OXH_ERR_ARG, "oxh_interaction_compute_forces: slot pools take their states through oxh_pool_admit");
	oxh::Batch &b = B->b;
	HIPCHK(hipSetDevice(b.device));
	hipStream_t st = stream ? (hipStream_t) stream : b.stream;
	// work queued on the batch's own stream comes first
	if(st != b.stream) HIPCHK(hipStreamSynchronize(b.stream));
	const int bs = 256, g = (b.N + bs - 1) / bs;
	k_load_caller_state<<<g, bs, 0, st>>>(b.N, (const double4 *) d_pos, (const double4 *) d_quat, b.orig, b.pinfo, b.bonds, b.d_sys,
			b.rlist, b.pos, b.q64, b.q32, b.bpos, b.back64, b.d_state, b.any_rebuild, b.blist);
	std::vector<oxh_sys_state> sst(b.nsys);
	for(int attempt = 0;; attempt++) {
		if(attempt == 12) return fail(OXH_ERR_SYSTEM, "neighbour list keeps overflowing");
		// lists of the systems whose particles moved more than the skin
		// (every system after a capacity change), then the forces
		oxh::launch_cells(st, b, b.pos);
		oxh::launch_build_list(st, b, 0, b.pos);
		if(B->persist) oxh::launch_persist_forces(st, b, false, B->persist_block);
		else oxh::launch_forces(st, b, false);
		HIPCHK(hipMemcpyAsync(sst.data(), b.d_state, sizeof(oxh_sys_state) * b.nsys, hipMemcpyDeviceToHost, st));
		HIPCHK(hipStreamSynchronize(st));
		HIPCHK(hipGetLastError());
		bool overflow = false;
		for(auto &x : sst) overflow |= (x.error & oxh::ERR_OVERFLOW) != 0;
		if(!overflow) {
			// the systems rebuilt count one list build each, as in a step
			k_count_rebuilds<<<(b.nsys + 63) / 64, 64, 0, st>>>(b.nsys, b.d_state, b.any_rebuild);
			break;
		}
		TRY(grow_lists(B, sst));
		HIPCHK(hipStreamSynchronize(b.stream));
		k_clear_need<<<(b.nsys + 63) / 64, 64, 0, st>>>(b.nsys, b.d_state, 1, b.any_rebuild);
	}
	if(b.fp64)
		k_add_to_caller<double4><<<g, bs, 0, st>>>(b.N, b.F64, b.T64, b.orig, b.pinfo, b.d_sys, (double4 *) d_forces, (double4 *) d_torques);
	else
		k_add_to_caller<float4><<<g, bs, 0, st>>>(b.N, b.F, b.T, b.orig, b.pinfo, b.d_sys, (float4 *) d_forces, (float4 *) d_torques);
	HIPCHK(hipGetLastError());
	return OXH_OK;
}

oxh_status oxh_batch_run(oxh_batch *B, long long n_steps) {
	if(!B) return fail(OXH_ERR_ARG, "null batch");
	if(n_steps < 0) return fail(OXH_ERR_ARG, "negative step count");
	oxh::Batch &b = B->b;
	HIPCHK(hipSetDevice(b.device));
	long long done = 0;
	int retries = 0;
	while(done < n_steps) {
		if(B->sort_interval > 0 && B->since_sort >= B->sort_interval && !B->persist && !B->pool) {
			if(oxh::launch_hilbert_sort(b.stream, b) != 0) return fail(OXH_ERR_ARG, "too many systems for the sort key");
			TRY(build_all(B));
			B->since_sort = 0;
		}
		long long todo = std::min<long long>(B->persist ? B->persist_chunk : B->chunk, n_steps - done);
		// chunks end where a sort or a re-boxing is due, so that these
		// happen at the same steps whatever the calls' lengths
		if(!B->persist && !B->pool && B->sort_interval > 0) todo = std::min(todo, B->sort_interval - B->since_sort);
		if(!B->pool && B->fix_every > 0) todo = std::min(todo, std::max(1LL, B->fix_every - B->since_fix));
		bool aligned = !B->any_bussi || !B->bussi_aligned || (b.step % B->bussi_ns == 0);
		if(B->pool) {
			// every slot advances by at most its remaining steps; slots
			// with nothing left stay frozen
			bool any = false;
			for(int s = 0; s < b.nsys; s++) {
				long long left = B->n[s] > 0 ? B->target[s] - B->done_steps[s] : 0;
				B->h_limit[s] = (int) std::max(0LL, std::min(left, todo));
				any |= B->h_limit[s] > 0;
			}
			if(!any) break;
			HIPCHK(hipMemcpyAsync(b.limit, B->h_limit.data(), b.nsys * sizeof(int), hipMemcpyHostToDevice, b.stream));
		}
		TRY(backup(B));
		if(B->persist) {
			oxh::launch_persist(b.stream, b, (int) todo, B->persist_block);
		}
		else if(todo == B->chunk && aligned && B->use_graph) {
			if(!b.graph_exec) {
				hipGraph_t g;
				HIPCHK(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
				enqueue_steps(B, B->chunk, 0);
				HIPCHK(hipStreamEndCapture(b.stream, &g));
				HIPCHK(hipGraphInstantiate(&b.graph_exec, g, nullptr, nullptr, 0));
				HIPCHK(hipGraphDestroy(g));
			}
			HIPCHK(hipGraphLaunch(b.graph_exec, b.stream));
		}
		else {
			enqueue_steps(B, (int) todo, b.step);
		}
		std::vector<oxh_sys_state> st;
		TRY(read_states(B, st));
		bool overflow = false;
		int bad = -1, bad_ext = -1;
		for(int s = 0; s < b.nsys; s++) {
			if(st[s].error & oxh::ERR_OVERFLOW) overflow = true;
			if((st[s].error & 1) && bad < 0) bad = s;
			if((st[s].error & oxh::ERR_EXT) && bad_ext < 0) bad_ext = s;
		}
		if(overflow) {
			// replay the chunk from its start with larger buffers and the
			// same lists
			if(++retries > 8) return fail(OXH_ERR_SYSTEM, "neighbour list keeps overflowing");
			TRY(restore(B));
			TRY(grow_lists(B, st));
			TRY(rebuild_from_rlist(B));
			continue;
		}
		b.step += todo;
		done += todo;
		B->since_sort += todo;
		B->since_fix += todo;
		if(B->fix_every > 0 && B->since_fix >= B->fix_every && !B->pool) {
			TRY(fix_diffusion_impl(B, nullptr));
			B->since_fix = 0;
		}
		if(B->pool) {
			for(int s = 0; s < b.nsys; s++) {
				B->done_steps[s] += B->h_limit[s];
				// a broken bond ends that job only: its slot stops here
				if((st[s].error & 1) && B->n[s] > 0) B->target[s] = B->done_steps[s];
			}
			continue;
		}
		if(bad >= 0) return fail(OXH_ERR_SYSTEM, "system %d: a bonded pair exceeded the FENE range", bad);
		if(bad_ext >= 0) return fail(OXH_ERR_SYSTEM, "system %d: a particle is beyond the radius of a yukawa_sphere force", bad_ext);
	}
	return OXH_OK;
}

oxh_status oxh_pool_create(int device, int n_slots, const int *capacity, const int *cell_budget, oxh_batch **out) {
	if(!out || n_slots <= 0 || !capacity) return fail(OXH_ERR_ARG, "oxh_pool_create: bad arguments");
	std::vector<int> cells(n_slots, 0);
	for(int s = 0; s < n_slots; s++) {
		if(capacity[s] <= 0) return fail(OXH_ERR_ARG, "slot %d: capacity must be > 0", s);
		// default budget: one cell per ~2 particles, enough for a system
		// of that size at any density
		cells[s] = cell_budget ? std::max(0, cell_budget[s]) : (capacity[s] > 1024 ? std::max(27, capacity[s] / 2) : 0);
	}
	return create_impl(device, n_slots, nullptr, capacity, cells.data(), out);
}

oxh_status oxh_pool_admit(oxh_batch *B, int slot, const oxh_system_desc *desc, long long n_steps) {
	TRY(check_sys(B, slot));
	if(!B->pool) return fail(OXH_ERR_ARG, "not a slot pool");
	if(!desc || n_steps < 0) return fail(OXH_ERR_ARG, "oxh_pool_admit: bad arguments");
	if(B->n[slot] > 0) return fail(OXH_ERR_ARG, "slot %d is in use", slot);
	const oxh_system_desc &d = *desc;
	if(d.N <= 0 || d.N > B->cap[slot]) return fail(OXH_ERR_ARG, "slot %d holds at most %d nucleotides (job has %d)", slot, B->cap[slot], d.N);
	if(d.backend_precision != 0) return fail(OXH_ERR_ARG, "slot pools run the mixed precision form only");
	if(d.interaction != OXH_INTERACTION_DNA2) return fail(OXH_ERR_ARG, "slot pools run the oxDNA2 model only");
	oxh::Batch &b = B->b;
	oxh_sys S;
	double rv = 0;
	TRY(sys_setup(B, d, slot, S, rv));
	B->bussi_aligned = false;
	const int f = B->first[slot];
	int max_occ = 1;
	double est = 16;
	sys_cells(d, S, rv, b.h_sys[slot].cell_first, B->cell_budget[slot], max_occ, est);
	if(S.use_cells && B->persist) return fail(OXH_ERR_ARG, "slot %d: the workgroup engine takes no cell lists", slot);
	S.far_back = (b.blist && S.use_cells && !S.small_box) ? 1 : 0;
	S.first = f;
	S.n = d.N;
	// host staging and upload of the slot's particles
	std::vector<double4> hpos(d.N), hvel(d.N), hL(d.N), hq(d.N);
	std::vector<int2> hbonds(d.N), hinfo(d.N);
	std::vector<int> horig(d.N);
	stage_system(d, slot, f, hpos.data(), hvel.data(), hL.data(), hq.data(), hbonds.data(), hinfo.data(), horig.data());
	oxh_sys_state hst;
	initial_state(d, hst);
	B->n_strands[slot] = strands_of(d, B->strand_of[slot]);
	hipStream_t st = b.stream;
	HIPCHK(hipStreamSynchronize(st));   // staging of an earlier admission may still be in flight
	HIPCHK(hipMemcpyAsync(b.pos + f, hpos.data(), d.N * sizeof(double4), hipMemcpyHostToDevice, st));
	HIPCHK(hipMemcpyAsync(b.rlist + f, hpos.data(), d.N * sizeof(double4), hipMemcpyHostToDevice, st));
	HIPCHK(hipMemcpyAsync(b.vel + f, hvel.data(), d.N * sizeof(double4), hipMemcpyHostToDevice, st));
	HIPCHK(hipMemcpyAsync(b.Lm + f, hL.data(), d.N * sizeof(double4), hipMemcpyHostToDevice, st));
	HIPCHK(hipMemcpyAsync(b.q64 + f, hq.data(), d.N * sizeof(double4), hipMemcpyHostToDevice, st));
	HIPCHK(hipMemcpyAsync(b.bonds + f, hbonds.data(), d.N * sizeof(int2), hipMemcpyHostToDevice, st));
	HIPCHK(hipMemcpyAsync(b.pinfo + f, hinfo.data(), d.N * sizeof(int2), hipMemcpyHostToDevice, st));
	HIPCHK(hipMemcpyAsync(b.orig + f, horig.data(), d.N * sizeof(int), hipMemcpyHostToDevice, st));
	HIPCHK(hipMemsetAsync(b.pstep + f, 0, d.N * sizeof(long long), st));
	HIPCHK(hipMemsetAsync(b.img + f, 0, d.N * sizeof(int4), st));
	HIPCHK(hipMemsetAsync(b.F + f, 0, d.N * sizeof(float4), st));
	HIPCHK(hipMemsetAsync(b.T + f, 0, d.N * sizeof(float4), st));
	b.h_sys[slot] = S;
	HIPCHK(hipMemcpyAsync(b.d_sys + slot, &b.h_sys[slot], sizeof(oxh_sys), hipMemcpyHostToDevice, st));
	HIPCHK(hipMemcpyAsync(b.d_state + slot, &hst, sizeof(hst), hipMemcpyHostToDevice, st));
	B->n[slot] = d.N;
	B->target[slot] = n_steps;
	B->done_steps[slot] = 0;
	if(d.refresh_vel) oxh::launch_refresh_vel(st, b, f, d.N);
	oxh::launch_prepare_mirror(st, b);
	oxh::launch_slot_of(st, b);
	// the slot's first list is built from rlist: its backbone sites are the
	// recorded ones
	if(S.far_back) HIPCHK(hipMemcpyAsync(b.blist + f, b.bpos + f, d.N * sizeof(float4), hipMemcpyDeviceToDevice, st));
	// reduction blocks follow the slots' sizes
	std::vector<int3> rb;
	std::vector<int2> rsb;
	red_layout(B, rb, rsb, b.nred);
	HIPCHK(hipMemcpyAsync(b.red_blocks, rb.data(), b.nred * sizeof(int3), hipMemcpyHostToDevice, st));
	HIPCHK(hipMemcpyAsync(b.red_sys_blocks, rsb.data(), b.nsys * sizeof(int2), hipMemcpyHostToDevice, st));
	// capacity for this system's cells and neighbours; growing rebuilds the
	// other slots' lists exactly (from their last build positions)
	bool regrow = false;
	if(S.use_cells && 3 * max_occ > b.cell_cap) {
		TRY(alloc_cell_slots(B, 3 * max_occ));
		regrow = true;
	}
	if((int) std::ceil(est) > b.max_neigh) {
		HIPCHK(hipStreamSynchronize(st));
		HIPCHK(hipFree(b.nbr));
		B->device_bytes -= (size_t) b.nstride * b.max_neigh * sizeof(int);
		b.max_neigh = std::min((int) std::ceil(est), std::max(b.N, 1));
		TRY(dalloc(B, &b.nbr, (size_t) b.nstride * b.max_neigh));
		regrow = true;
	}
	if(regrow && b.graph_exec) {
		(void) hipGraphExecDestroy(b.graph_exec);
		b.graph_exec = nullptr;
	}
	// the new system's list (its rlist is its position); with regrown
	// buffers every slot's list is rebuilt from its rlist, which gives the
	// same lists as before
	TRY(rebuild_from_rlist(B, regrow ? -1 : slot));
	enqueue_forces(B);
	return sync_and_check(B);
}

oxh_status oxh_pool_release(oxh_batch *B, int slot) {
	TRY(check_sys(B, slot));
	if(!B->pool) return fail(OXH_ERR_ARG, "not a slot pool");
	oxh::Batch &b = B->b;
	HIPCHK(hipStreamSynchronize(b.stream));
	oxh_sys &S = b.h_sys[slot];
	S.n = 0;
	S.use_cells = 0;
	S.ncell = 0;
	B->n[slot] = 0;
	B->target[slot] = B->done_steps[slot] = 0;
	oxh_sys_state hst;
	std::memset(&hst, 0, sizeof(hst));
	HIPCHK(hipMemcpyAsync(b.d_sys + slot, &S, sizeof(oxh_sys), hipMemcpyHostToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(b.d_state + slot, &hst, sizeof(hst), hipMemcpyHostToDevice, b.stream));
	std::vector<int3> rb;
	std::vector<int2> rsb;
	red_layout(B, rb, rsb, b.nred);
	HIPCHK(hipMemcpyAsync(b.red_blocks, rb.data(), b.nred * sizeof(int3), hipMemcpyHostToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(b.red_sys_blocks, rsb.data(), b.nsys * sizeof(int2), hipMemcpyHostToDevice, b.stream));
	return sync_and_check(B);
}

oxh_status oxh_pool_status(oxh_batch *B, long long *steps_done, long long *steps_target, int *n_nucleotides, int *error) {
	if(!B || !B->pool) return fail(OXH_ERR_ARG, "not a slot pool");
	std::vector<oxh_sys_state> st;
	TRY(read_states(B, st));
	for(int s = 0; s < B->b.nsys; s++) {
		if(steps_done) steps_done[s] = B->done_steps[s];
		if(steps_target) steps_target[s] = B->target[s];
		if(n_nucleotides) n_nucleotides[s] = B->n[s];
		if(error) error[s] = B->n[s] > 0 ? st[s].error : 0;
	}
	return OXH_OK;
}

oxh_status oxh_batch_hilbert_sort(oxh_batch *B) {
	if(!B) return fail(OXH_ERR_ARG, "null batch");
	HIPCHK(hipSetDevice(B->b.device));
	if(oxh::launch_hilbert_sort(B->b.stream, B->b) != 0) return fail(OXH_ERR_ARG, "too many systems for the sort key");
	TRY(sync_and_check(B));
	return oxh_batch_update_lists(B, 1);
}

oxh_status oxh_batch_fix_diffusion(oxh_batch *B, int *n_moved) {
	if(!B) return fail(OXH_ERR_ARG, "null batch");
	HIPCHK(hipSetDevice(B->b.device));
	return fix_diffusion_impl(B, n_moved);
}

// intervals of the host-side events are whole chunks, so that they fall on
// the same steps however a run is split into oxh_batch_run calls
static long long round_to_chunks(const oxh_batch *B, long long steps) {
	const long long c = std::max(1, B->chunk);
	return steps > 0 ? (steps + c - 1) / c * c : 0;
}

oxh_status oxh_batch_set_fix_diffusion(oxh_batch *B, long long every) {
	if(!B || every < 0) return fail(OXH_ERR_ARG, "bad arguments");
	B->fix_every = round_to_chunks(B, every);
	B->since_fix = 0;
	return OXH_OK;
}

oxh_status oxh_batch_set_sort_interval(oxh_batch *B, long long steps) {
	if(!B || steps < 0) return fail(OXH_ERR_ARG, "bad arguments");
	B->sort_interval = round_to_chunks(B, steps);
	B->since_sort = B->sort_interval;   // sort before the next chunk
	return OXH_OK;
}

oxh_status oxh_batch_sync(oxh_batch *B) {
	if(!B) return fail(OXH_ERR_ARG, "null batch");
	HIPCHK(hipSetDevice(B->b.device));
	return sync_and_check(B);
}


oxh_status oxh_batch_get_state(oxh_batch *B, int sys, double *pos, double *a1, double *a3, double *vel, double *L) {
	TRY(check_sys(B, sys));
	int f = B->first[sys], n = B->n[sys];
	std::vector<double4> p, q, v, l;
	std::vector<int> o;
	TRY(download(B, p, B->b.pos, f, n));
	TRY(download(B, q, B->b.q64, f, n));
	TRY(download(B, v, B->b.vel, f, n));
	TRY(download(B, l, B->b.Lm, f, n));
	TRY(download(B, o, B->b.orig, f, n));
	TRY(sync_and_check(B));
	for(int k = 0; k < n; k++) {
		int i = o[k];
		if(pos) {
			pos[3 * i] = p[k].x;
			pos[3 * i + 1] = p[k].y;
			pos[3 * i + 2] = p[k].z;
		}
		double x1[3], x3[3];
		quat_to_axes(q[k], x1, x3);
		for(int c = 0; c < 3; c++) {
			if(a1) a1[3 * i + c] = x1[c];
			if(a3) a3[3 * i + c] = x3[c];
		}
		if(vel) {
			vel[3 * i] = v[k].x;
			vel[3 * i + 1] = v[k].y;
			vel[3 * i + 2] = v[k].z;
		}
		if(L) {
			L[3 * i] = l[k].x;
			L[3 * i + 1] = l[k].y;
			L[3 * i + 2] = l[k].z;
		}
	}
	return OXH_OK;
}

oxh_status oxh_batch_set_external_forces(oxh_batch *B, int sys, int n, const oxh_ext_force *forces, long long step0) {
	TRY(check_sys(B, sys));
	if(B->pool) return fail(OXH_ERR_ARG, "external forces are set on batches, not slot pools");
	if(n < 0 || (n > 0 && !forces)) return fail(OXH_ERR_ARG, "oxh_batch_set_external_forces: bad arguments");
	oxh::Batch &b = B->b;
	const int N = B->n[sys];
	std::vector<std::pair<int, oxh_ext_rec>> recs;
	std::vector<std::pair<std::vector<int>, std::vector<int>>> groups;
	for(int k = 0; k < n; k++) {
		const oxh_ext_force &f = forces[k];
		if(f.particle < 0 || f.particle >= N) return fail(OXH_ERR_ARG, "external force %d: no particle %d in system %d", k, f.particle, sys);
		oxh_ext_rec r;
		std::memset(&r, 0, sizeof(r));
		r.type = f.type;
		// LR_vector::normalize: times the reciprocal of the module
		// (src/Utilities/LR_vector.h:86-89,131-136)
		auto unit = [](const double *v, double *out) -> bool {
			const double m = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
			if(!(m > 0)) return false;
			const double inv = 1. / m;
			for(int c = 0; c < 3; c++) out[c] = v[c] * inv;
			return true;
		};
		auto copy3 = [](const double *v, double *out) {
			for(int c = 0; c < 3; c++) out[c] = v[c];
		};
		auto in_system = [&](int i) { return i >= 0 && i < N; };
		switch(f.type) {
		case OXH_EXT_STRING:
			r.a = f.F0;
			r.b = f.rate;
			if(f.dir_as_centre) {
				r.flags = 2;
				copy3(f.dir, r.v1);
			}
			else if(!unit(f.dir, r.v1)) return fail(OXH_ERR_ARG, "external force %d: zero dir", k);
			break;
		case OXH_EXT_TRAP:
		case OXH_EXT_LOWDIM_TRAP:
			r.a = f.stiff;
			r.b = f.rate;
			if(!unit(f.dir, r.v1)) return fail(OXH_ERR_ARG, "external force %d: zero dir", k);
			copy3(f.pos0, r.v2);
			if(f.type == OXH_EXT_LOWDIM_TRAP) r.flags = (f.visibility[0] ? 1 : 0) | (f.visibility[1] ? 2 : 0) | (f.visibility[2] ? 4 : 0);
			break;
		case OXH_EXT_MUTUAL_TRAP:
		case OXH_EXT_CONSTANT_TRAP:
			if(!in_system(f.ref_particle))
				return fail(OXH_ERR_ARG, "external force %d: no reference particle %d in system %d", k, f.ref_particle, sys);
			r.ref = f.ref_particle;
			r.flags = f.pbc ? 1 : 0;
			r.a = f.stiff;
			r.b = f.type == OXH_EXT_MUTUAL_TRAP ? f.stiff_rate : 0.0;
			r.c = f.r0;
			r.d = f.type == OXH_EXT_MUTUAL_TRAP ? f.rate : 0.0;
			break;
		case OXH_EXT_REPULSION_PLANE:
			r.a = f.stiff;
			r.b = f.position;
			r.c = f.end_position;
			r.d = f.v;
			if(!unit(f.dir, r.v1)) return fail(OXH_ERR_ARG, "external force %d: zero dir", k);
			break;
		case OXH_EXT_ATTRACTION_PLANE:
			r.a = f.stiff;
			r.b = f.position;
			if(!unit(f.dir, r.v1)) return fail(OXH_ERR_ARG, "external force %d: zero dir", k);
			break;
		case OXH_EXT_PLANE_MOVING:
			if(!in_system(f.ref_particle) || !in_system(f.ref_last) || f.ref_last < f.ref_particle)
				return fail(OXH_ERR_ARG, "external force %d: reference particles %d..%d are not a range of system %d", k, f.ref_particle,
						f.ref_last, sys);
			r.a = f.stiff;
			r.ref = f.ref_particle;
			r.ref2 = f.ref_last;
			if(!unit(f.dir, r.v1)) return fail(OXH_ERR_ARG, "external force %d: zero dir", k);
			break;
		case OXH_EXT_SPHERE:
			r.a = f.stiff;
			r.b = f.rate;
			r.c = f.r0;
			r.d = f.r_ext;
			copy3(f.center, r.v1);
			break;
		case OXH_EXT_SPHERE_SMOOTH:
			r.a = f.stiff;
			r.c = f.r0;
			r.d = f.r_ext;
			r.e = f.smooth;
			r.f = f.alpha;
			copy3(f.center, r.v1);
			break;
		case OXH_EXT_SPHERE_MOVING:
			r.a = f.stiff;
			r.b = f.rate;
			r.c = f.r0;
			r.d = f.r_ext;
			r.steps = f.steps;
			copy3(f.pos0, r.v1);
			copy3(f.target, r.v2);
			break;
		case OXH_EXT_KEPLER_POINSOT:
			if(!(f.apex > f.base)) return fail(OXH_ERR_ARG, "external force %d: repulsive_kepler_poinsot needs apex > base", k);
			if(!(f.base_radius > 0)) return fail(OXH_ERR_ARG, "external force %d: repulsive_kepler_poinsot needs base_radius > 0", k);
			if(!(f.kappa > 0)) return fail(OXH_ERR_ARG, "external force %d: repulsive_kepler_poinsot needs kappa > 0", k);
			r.a = f.stiff;
			r.b = f.rate;
			r.c = f.apex;
			r.d = f.base;
			r.e = f.base_radius;
			r.f = f.kappa;
			copy3(f.center, r.v1);
			break;
		case OXH_EXT_LJ_WALL:
		case OXH_EXT_LJ_CONE:
			if(f.n <= 0 || f.n % 2 != 0) return fail(OXH_ERR_ARG, "external force %d: n (%d) must be a positive even integer", k, f.n);
			if(!(f.sigma > 0)) return fail(OXH_ERR_ARG, "external force %d: sigma must be > 0", k);
			r.a = f.stiff;
			r.c = f.sigma;
			r.d = f.cutoff;
			r.n = f.n;
			if(!unit(f.dir, r.v1)) return fail(OXH_ERR_ARG, "external force %d: zero dir", k);
			if(f.type == OXH_EXT_LJ_WALL) r.b = f.position;
			else {
				copy3(f.pos0, r.v2);
				r.e = std::sin(f.alpha);
				r.f = std::cos(f.alpha);
				r.g = std::tan(f.alpha);
			}
			break;
		case OXH_EXT_TWIST:
			r.a = f.stiff;
			r.b = f.rate;
			r.c = f.F0;
			if(!unit(f.axis, r.v1)) return fail(OXH_ERR_ARG, "external force %d: zero axis", k);
			copy3(f.pos0, r.v2);
			copy3(f.center, r.v3);
			copy3(f.mask, r.v4);
			break;
		case OXH_EXT_CENTRAL:
			r.a = f.F0;
			r.c = f.inner_cut_off;
			r.d = f.outer_cut_off;
			copy3(f.center, r.v1);
			break;
		case OXH_EXT_ELLIPSOID:
			r.a = f.stiff;
			copy3(f.center, r.v1);
			copy3(f.r_1, r.v2);
			copy3(f.r_2, r.v3);
			break;
		case OXH_EXT_COM: {
			if(f.n_com <= 0 || f.n_ref <= 0 || !f.com_list || !f.ref_list)
				return fail(OXH_ERR_ARG, "external force %d: com needs non-empty com_list and ref_list", k);
			// the reference keeps both lists as sets of particles
			std::vector<int> com(f.com_list, f.com_list + f.n_com), ref(f.ref_list, f.ref_list + f.n_ref);
			for(auto *v : {&com, &ref}) {
				std::sort(v->begin(), v->end());
				v->erase(std::unique(v->begin(), v->end()), v->end());
				for(int i : *v)
					if(!in_system(i)) return fail(OXH_ERR_ARG, "external force %d: no particle %d in system %d", k, i, sys);
			}
			int g = 0;
			while(g < (int) groups.size() && !(groups[g].first == com && groups[g].second == ref)) g++;
			if(g == (int) groups.size()) groups.push_back({com, ref});
			r.a = f.stiff;
			r.b = f.rate;
			r.c = f.r0;
			r.ref = g;
			r.n = (int) com.size();
			break;
		}
		case OXH_EXT_YUKAWA_SPHERE:
			if(f.n <= 0) return fail(OXH_ERR_ARG, "external force %d: WCA_n must be > 0", k);
			if(!(f.debye_length > 0)) return fail(OXH_ERR_ARG, "external force %d: debye_length must be > 0", k);
			r.a = f.radius;
			r.b = f.epsilon;
			r.c = f.sigma;
			r.d = f.debye_length;
			r.e = f.debye_A;
			r.f = f.cutoff;
			r.g = f.sigma * std::pow(2., 1. / f.n);
			r.n = f.n;
			copy3(f.center, r.v1);
			break;
		default: return fail(OXH_ERR_ARG, "external force %d: unknown type %d", k, f.type);
		}
		recs.push_back({f.particle, r});
	}
	HIPCHK(hipStreamSynchronize(b.stream));
	if(B->ext_of.empty()) B->ext_of.resize(b.nsys);
	if(B->grp_of.empty()) B->grp_of.resize(b.nsys);
	std::stable_sort(recs.begin(), recs.end(), [](const std::pair<int, oxh_ext_rec> &x, const std::pair<int, oxh_ext_rec> &y) {
		return x.first < y.first;
	});
	B->ext_of[sys] = recs;
	B->grp_of[sys] = groups;
	// CSR over every system's particles: offsets per (first + particle);
	// com groups numbered over the batch, systems in order
	std::vector<int> off(b.N + 1, 0), gbase(b.nsys + 1, 0), idx, gsys;
	std::vector<int4> grp;
	std::vector<oxh_ext_rec> all;
	for(int s = 0; s < b.nsys; s++) {
		gbase[s + 1] = gbase[s] + (int) B->grp_of[s].size();
		for(auto &x : B->ext_of[s]) off[B->first[s] + x.first + 1]++;
		for(auto &gr : B->grp_of[s]) {
			const int c0 = (int) idx.size();
			idx.insert(idx.end(), gr.first.begin(), gr.first.end());
			const int r0 = (int) idx.size();
			idx.insert(idx.end(), gr.second.begin(), gr.second.end());
			grp.push_back(make_int4(c0, (int) gr.first.size(), r0, (int) gr.second.size()));
			gsys.push_back(s);
		}
	}
	for(int i = 0; i < b.N; i++) off[i + 1] += off[i];
	all.reserve(off[b.N]);
	for(int s = 0; s < b.nsys; s++)
		for(auto &x : B->ext_of[s]) {   // systems in order, particles sorted inside each
			all.push_back(x.second);
			if(x.second.type == OXH_EXT_COM) all.back().ref += gbase[s];
		}
	void *old[] = {b.ext_off, b.ext_rec, b.ext_grp, b.ext_grp_sys, b.ext_idx, b.ext_com};
	for(void *p : old)
		if(p) HIPCHK(hipFree(p));
	b.ext_off = nullptr;
	b.ext_rec = nullptr;
	b.ext_grp = nullptr;
	b.ext_grp_sys = nullptr;
	b.ext_idx = nullptr;
	b.ext_com = nullptr;
	b.n_ext = (int) all.size();
	b.n_grp = (int) grp.size();
	if(b.n_ext > 0) {
		TRY(dalloc(B, &b.ext_off, b.N + 1));
		TRY(dalloc(B, &b.ext_rec, all.size()));
		HIPCHK(hipMemcpyAsync(b.ext_off, off.data(), (b.N + 1) * sizeof(int), hipMemcpyHostToDevice, b.stream));
		HIPCHK(hipMemcpyAsync(b.ext_rec, all.data(), all.size() * sizeof(oxh_ext_rec), hipMemcpyHostToDevice, b.stream));
	}
	if(b.n_grp > 0) {
		TRY(dalloc(B, &b.ext_grp, grp.size()));
		TRY(dalloc(B, &b.ext_grp_sys, gsys.size()));
		TRY(dalloc(B, &b.ext_idx, idx.size()));
		TRY(dalloc(B, &b.ext_com, 6 * grp.size()));
		HIPCHK(hipMemcpyAsync(b.ext_grp, grp.data(), grp.size() * sizeof(int4), hipMemcpyHostToDevice, b.stream));
		HIPCHK(hipMemcpyAsync(b.ext_grp_sys, gsys.data(), gsys.size() * sizeof(int), hipMemcpyHostToDevice, b.stream));
		HIPCHK(hipMemcpyAsync(b.ext_idx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, b.stream));
	}
	oxh_sys &S = b.h_sys[sys];
	S.has_ext = recs.empty() ? 0 : 1;
	// the kernels count a particle's steps in pstep (the step being
	// completed in a step's force phase is pstep - 1): after `done` steps
	// the next force evaluation is pstep - 1 = done, which must read step0
	std::vector<long long> ps;
	TRY(download(B, ps, b.pstep, B->first[sys], 1));
	TRY(sync_and_check(B));
	const long long done = ps.empty() ? 0 : ps[0];
	S.ext_step0 = step0 - done;
	HIPCHK(hipMemcpyAsync(b.d_sys + sys, &S, sizeof(oxh_sys), hipMemcpyHostToDevice, b.stream));
	// the workgroup engine has no external-force phase: such batches step
	// with the grid engine
	if(b.n_ext > 0 && B->persist) B->persist = false;
	if(b.graph_exec) {
		(void) hipGraphExecDestroy(b.graph_exec);
		b.graph_exec = nullptr;
	}
	k_clear_err<<<(b.nsys + 63) / 64, 64, 0, b.stream>>>(b.nsys, b.d_state, oxh::ERR_EXT);
	enqueue_forces(B);
	std::vector<oxh_sys_state> st;
	TRY(read_states(B, st));
	if(st[sys].error & oxh::ERR_EXT) return fail(OXH_ERR_SYSTEM, "system %d: a particle is beyond the radius of a yukawa_sphere force", sys);
	return OXH_OK;
}

oxh_status oxh_batch_get_images(oxh_batch *B, int sys, int *img3) {
	TRY(check_sys(B, sys));
	if(!img3) return fail(OXH_ERR_ARG, "null image buffer");
	const int f = B->first[sys], n = B->n[sys];
	std::vector<int4> m;
	std::vector<int> o;
	TRY(download(B, m, B->b.img, f, n));
	TRY(download(B, o, B->b.orig, f, n));
	TRY(sync_and_check(B));
	for(int k = 0; k < n; k++) {
		img3[3 * o[k]] = m[k].x;
		img3[3 * o[k] + 1] = m[k].y;
		img3[3 * o[k] + 2] = m[k].z;
	}
	return OXH_OK;
}

oxh_status oxh_batch_set_images(oxh_batch *B, int sys, const int *img3) {
	TRY(check_sys(B, sys));
	if(!img3) return fail(OXH_ERR_ARG, "null image buffer");
	const int f = B->first[sys], n = B->n[sys];
	std::vector<int> o;
	TRY(download(B, o, B->b.orig, f, n));
	TRY(sync_and_check(B));
	std::vector<int4> m(n);
	for(int k = 0; k < n; k++) m[k] = make_int4(img3[3 * o[k]], img3[3 * o[k] + 1], img3[3 * o[k] + 2], 0);
	HIPCHK(hipMemcpyAsync(B->b.img + f, m.data(), n * sizeof(int4), hipMemcpyHostToDevice, B->b.stream));
	// forces that depend on absolute positions follow the new images
	if(B->b.n_ext > 0) enqueue_forces(B);
	return sync_and_check(B);
}

oxh_status oxh_batch_get_order(oxh_batch *B, int sys, int *order) {
	TRY(check_sys(B, sys));
	if(!order) return fail(OXH_ERR_ARG, "null order buffer");
	std::vector<int> o;
	TRY(download(B, o, B->b.orig, B->first[sys], B->n[sys]));
	TRY(sync_and_check(B));
	std::copy(o.begin(), o.end(), order);
	return OXH_OK;
}

oxh_status oxh_batch_get_forces(oxh_batch *B, int sys, double *F, double *T, double *Epart) {
	TRY(check_sys(B, sys));
	int f = B->first[sys], n = B->n[sys];
	std::vector<double4> hf(n), ht(n);
	std::vector<int> o;
	if(B->b.fp64) {
		TRY(download(B, hf, B->b.F64, f, n));
		TRY(download(B, ht, B->b.T64, f, n));
	}
	else {
		std::vector<float4> sf, sT;
		TRY(download(B, sf, B->b.F, f, n));
		TRY(download(B, sT, B->b.T, f, n));
		TRY(sync_and_check(B));
		for(int k = 0; k < n; k++) {
			hf[k] = make_double4(sf[k].x, sf[k].y, sf[k].z, sf[k].w);
			ht[k] = make_double4(sT[k].x, sT[k].y, sT[k].z, sT[k].w);
		}
	}
	TRY(download(B, o, B->b.orig, f, n));
	TRY(sync_and_check(B));
	for(int k = 0; k < n; k++) {
		int i = o[k];
		if(F) {
			F[3 * i] = hf[k].x;
			F[3 * i + 1] = hf[k].y;
			F[3 * i + 2] = hf[k].z;
		}
		if(T) {
			T[3 * i] = ht[k].x;
			T[3 * i + 1] = ht[k].y;
			T[3 * i + 2] = ht[k].z;
		}
		if(Epart) Epart[i] = hf[k].w;
	}
	return OXH_OK;
}

oxh_status oxh_batch_pair_quantities(oxh_batch *B, int sys, int kind, int n, const int *pairs, double *out) {
	TRY(check_sys(B, sys));
	if(kind < OXH_PAIR_HB_ENERGY || kind > OXH_PAIR_CENTRE_DISTANCE) return fail(OXH_ERR_ARG, "unknown pair quantity %d", kind);
	if(n < 0 || (n > 0 && (!pairs || !out))) return fail(OXH_ERR_ARG, "oxh_batch_pair_quantities: bad arguments");
	if(n == 0) return OXH_OK;
	oxh::Batch &b = B->b;
	if(b.model && kind == OXH_PAIR_NEAR_HB)
		return fail(OXH_ERR_ARG, "oxh_batch_pair_quantities: the near-HB criterion is not available for the oxRNA models");
	const int ns = B->n[sys];
	const std::vector<int> &sid = B->strand_of[sys];
	std::vector<int4> h(n);
	for(int k = 0; k < n; k++) {
		const int p = pairs[2 * k], q = pairs[2 * k + 1];
		if(p < 0 || q < 0 || p >= ns || q >= ns) return fail(OXH_ERR_ARG, "pair %d: particle index out of range", k);
		h[k] = make_int4(p, q, sid[p] == sid[q] ? 1 : 0, 0);
	}
	HIPCHK(hipSetDevice(b.device));
	int4 *d_pairs = nullptr;
	double *d_out = nullptr;
	HIPCHK(hipMalloc((void **) &d_pairs, (size_t) n * sizeof(int4)));
	if(hipMalloc((void **) &d_out, (size_t) n * sizeof(double)) != hipSuccess) {
		(void) hipFree(d_pairs);
		return fail(OXH_ERR_ALLOC, "oxh_batch_pair_quantities: allocation failed");
	}
	hipError_t e = hipMemcpyAsync(d_pairs, h.data(), (size_t) n * sizeof(int4), hipMemcpyHostToDevice, b.stream);
	if(e == hipSuccess) {
		oxh::launch_pair_quantities(b.stream, b, sys, kind, n, d_pairs, d_out);
		e = hipGetLastError();
	}
	if(e == hipSuccess) e = hipMemcpyAsync(out, d_out, (size_t) n * sizeof(double), hipMemcpyDeviceToHost, b.stream);
	if(e == hipSuccess) e = hipStreamSynchronize(b.stream);
	(void) hipFree(d_pairs);
	(void) hipFree(d_out);
	if(e != hipSuccess) return fail(OXH_ERR_HIP, "oxh_batch_pair_quantities: %s", hipGetErrorString(e));
	return OXH_OK;
}

}  // extern "C"

namespace {

// potential energy of system sys at the current state (the batch's force
// evaluation, then its FP64 reduction)
oxh_status system_energy(oxh_batch *B, int sys, double &U) {
	enqueue_forces(B);
	oxh::launch_reduce(B->b.stream, B->b);
	std::vector<double> r((size_t) B->b.nsys * oxh::RED_W);
	HIPCHK(hipMemcpyAsync(r.data(), B->b.red_out, r.size() * sizeof(double), hipMemcpyDeviceToHost, B->b.stream));
	TRY(sync_and_check(B));
	U = r[(size_t) sys * oxh::RED_W];
	return OXH_OK;
}

// the box of system sys set to L, with everything that follows from it:
// inverse sides, FP32 copies, the FP32 rounding bound of the list
// predicate, the small-box flag and the cells (the cell count per side only
// goes down, so the system keeps its cell range; cells stay at least as
// wide as its stencil needs)
oxh_status set_system_box(oxh_batch *B, int sys, const double L[3]) {
	oxh_sys &S = B->b.h_sys[sys];
	const double rv = std::sqrt(S.rv2);
	for(int k = 0; k < 3; k++) {
		if(!(L[k] > 0)) return fail(OXH_ERR_ARG, "system %d: box side %d must be > 0", sys, k);
		S.box[k] = L[k];
		S.inv_box[k] = 1.0 / L[k];
		S.boxf[k] = (float) L[k];
		S.inv_boxf[k] = (float) (1.0 / L[k]);
	}
	const double lmax = std::max(L[0], std::max(L[1], L[2]));
	const double dr = 4.0 * lmax * std::ldexp(1.0, -24);
	S.eps2 = (float) (4.0 * (2.0 * rv * dr * std::sqrt(3.0) + 3.0 * dr * dr + 8.0 * rv * rv * std::ldexp(1.0, -24)));
	S.small_box = 0;
	for(int k = 0; k < 3; k++)
		if(L[k] < 2.0 * (rv + 2.0 * 0.4814)) S.small_box = 1;
	// the lists are rebuilt after a box change, so the far-pair filter can
	// change with it: off in a small box, its margin follows the box
	if(S.small_box) S.far_back = 0;
	far_back_bounds(S, std::sqrt(S.skin2));
	if(S.use_cells) {
		for(int k = 0; k < 3; k++) {
			const int fit = (int) std::floor(L[k] * S.stencil / rv);
			S.nc[k] = std::min(S.nc[k], fit);
			if(S.nc[k] < 2 * S.stencil + 1) return fail(OXH_ERR_SYSTEM, "system %d: box side %d too short for its cells", sys, k);
			S.cell_w[k] = (float) (L[k] / S.nc[k]);
		}
		S.ncell = S.nc[0] * S.nc[1] * S.nc[2];
	}
	HIPCHK(hipMemcpyAsync(B->b.d_sys + sys, &S, sizeof(oxh_sys), hipMemcpyHostToDevice, B->b.stream));
	return sync_and_check(B);
}

}  // namespace

extern "C" {

oxh_status oxh_batch_energies(oxh_batch *B, double *out4) {
	if(!B || !out4) return fail(OXH_ERR_ARG, "bad arguments");
	HIPCHK(hipSetDevice(B->b.device));
	oxh::launch_reduce(B->b.stream, B->b);
	std::vector<double> r((size_t) B->b.nsys * oxh::RED_W);
	HIPCHK(hipMemcpyAsync(r.data(), B->b.red_out, r.size() * sizeof(double), hipMemcpyDeviceToHost, B->b.stream));
	TRY(sync_and_check(B));
	for(int s = 0; s < B->b.nsys; s++)
		for(int k = 0; k < 4; k++) out4[4 * s + k] = r[(size_t) s * oxh::RED_W + k];
	return OXH_OK;
}

oxh_status oxh_batch_energy_split(oxh_batch *B, int sys, double *E8) {
	TRY(check_sys(B, sys));
	enqueue_forces(B, true);
	int f = B->first[sys], n = B->n[sys];
	std::vector<double> e((size_t) 8 * n);
	for(int t = 0; t < 8; t++)
		HIPCHK(hipMemcpyAsync(e.data() + (size_t) t * n, B->b.esplit + (size_t) t * B->b.N + f, n * sizeof(double), hipMemcpyDeviceToHost,
				B->b.stream));
	TRY(sync_and_check(B));
	for(int t = 0; t < 8; t++) {
		double acc = 0;
		for(int k = 0; k < n; k++) acc += e[(size_t) t * n + k];
		E8[t] = acc;
	}
	return OXH_OK;
}

oxh_status oxh_batch_stress(oxh_batch *B, int sys, double *virial6, double *kinetic6) {
	TRY(check_sys(B, sys));
	enqueue_forces(B, true);
	const int f = B->first[sys], n = B->n[sys];
	std::vector<double> w((size_t) 6 * std::max(n, 1));
	for(int c = 0; c < 6; c++)
		HIPCHK(hipMemcpyAsync(w.data() + (size_t) c * n, B->b.stress + (size_t) c * B->b.N + f, n * sizeof(double), hipMemcpyDeviceToHost,
				B->b.stream));
	std::vector<double4> v;
	TRY(download(B, v, B->b.vel, f, n));
	TRY(sync_and_check(B));
	for(int c = 0; c < 6; c++) {
		double acc = 0.;
		for(int k = 0; k < n; k++) acc += w[(size_t) c * n + k];
		if(virial6) virial6[c] = acc;
	}
	if(kinetic6) {
		for(int c = 0; c < 6; c++) kinetic6[c] = 0.;
		for(const double4 &x : v) {
			kinetic6[0] += x.x * x.x;
			kinetic6[1] += x.y * x.y;
			kinetic6[2] += x.z * x.z;
			kinetic6[3] += x.x * x.y;
			kinetic6[4] += x.x * x.z;
			kinetic6[5] += x.y * x.z;
		}
	}
	return OXH_OK;
}

oxh_status oxh_batch_hb_count(oxh_batch *B, int *out) {
	if(!B || !out) return fail(OXH_ERR_ARG, "bad arguments");
	HIPCHK(hipSetDevice(B->b.device));
	B->b.want_hbcount = true;
	enqueue_forces(B);
	B->b.want_hbcount = false;
	std::vector<int> h(B->b.N);
	HIPCHK(hipMemcpyAsync(h.data(), B->b.hbcount, B->b.N * sizeof(int), hipMemcpyDeviceToHost, B->b.stream));
	TRY(sync_and_check(B));
	for(int s = 0; s < B->b.nsys; s++) {
		long long acc = 0;
		for(int i = B->first[s]; i < B->first[s] + B->n[s]; i++) acc += h[i];
		out[s] = (int) (acc / 2);
	}
	return OXH_OK;
}

oxh_status oxh_batch_get_pairs(oxh_batch *B, int sys, int *pairs, long long max, long long *n_out) {
	TRY(check_sys(B, sys));
	int f = B->first[sys], n = B->n[sys];
	const int M = B->b.max_neigh;
	std::vector<int2> nn;
	std::vector<int> all_orig;
	TRY(download(B, nn, B->b.nnbr, f, n));
	TRY(download(B, all_orig, B->b.orig, 0, B->b.N));
	std::vector<int> rows((size_t) M * n);
	HIPCHK(hipMemcpyAsync(rows.data(), B->b.nbr + (size_t) f * M, rows.size() * sizeof(int), hipMemcpyDeviceToHost, B->b.stream));
	TRY(sync_and_check(B));
	std::vector<std::pair<int, int>> pr;
	for(int k = 0; k < n; k++) {
		int i = all_orig[f + k];
		const int *row = rows.data() + (size_t) k * M;
		auto add = [&](int jb) {
			int j = all_orig[jb];
			pr.push_back(i < j ? std::make_pair(i, j) : std::make_pair(j, i));
		};
		for(int m = 0; m < nn[k].x; m++) add(row[m]);
		for(int m = 0; m < nn[k].y; m++) add(row[M - 1 - m]);
	}
	std::sort(pr.begin(), pr.end());
	pr.erase(std::unique(pr.begin(), pr.end()), pr.end());
	*n_out = (long long) pr.size();
	for(long long k = 0; k < (long long) pr.size() && k < max; k++) {
		pairs[2 * k] = pr[k].first;
		pairs[2 * k + 1] = pr[k].second;
	}
	return OXH_OK;
}

oxh_status oxh_batch_system_info(oxh_batch *B, int sys, long long *n_rebuilds, int *error) {
	TRY(check_sys(B, sys));
	std::vector<oxh_sys_state> st;
	TRY(read_states(B, st));
	if(n_rebuilds) *n_rebuilds = st[sys].n_rebuilds;
	if(error) *error = st[sys].error;
	return OXH_OK;
}

oxh_status oxh_batch_set_state(oxh_batch *B, int sys, const double *pos, const double *a1, const double *a3, const double *vel,
		const double *L) {
	TRY(check_sys(B, sys));
	oxh::Batch &b = B->b;
	int f = B->first[sys], n = B->n[sys];
	std::vector<int> o;
	std::vector<double4> hp, hq, hv, hl;
	TRY(download(B, o, b.orig, f, n));
	TRY(download(B, hp, b.pos, f, n));
	TRY(download(B, hq, b.q64, f, n));
	TRY(download(B, hv, b.vel, f, n));
	TRY(download(B, hl, b.Lm, f, n));
	TRY(sync_and_check(B));
	for(int k = 0; k < n; k++) {
		int i = o[k];
		if(pos) hp[k] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.0);
		if(a1 && a3) {
			double q[4];
			frame_to_quat(a1 + 3 * i, a3 + 3 * i, q);
			hq[k] = make_double4(q[0], q[1], q[2], q[3]);
		}
		if(vel) hv[k] = make_double4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0.0);
		if(L) hl[k] = make_double4(L[3 * i], L[3 * i + 1], L[3 * i + 2], 0.0);
	}
	HIPCHK(hipMemcpyAsync(b.pos + f, hp.data(), n * sizeof(double4), hipMemcpyHostToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(b.q64 + f, hq.data(), n * sizeof(double4), hipMemcpyHostToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(b.vel + f, hv.data(), n * sizeof(double4), hipMemcpyHostToDevice, b.stream));
	HIPCHK(hipMemcpyAsync(b.Lm + f, hl.data(), n * sizeof(double4), hipMemcpyHostToDevice, b.stream));
	oxh::launch_prepare_mirror(b.stream, b);
	TRY(sync_and_check(B));
	TRY(oxh_batch_update_lists(B, 1));
	return oxh_batch_compute_forces(B);
}

oxh_status oxh_batch_barostat_move(oxh_batch *B, int sys, double P, double delta_L, int isotropic, int molecular, const double *u4,
		int *accepted, double *dE) {
	TRY(check_sys(B, sys));
	if(B->pool) return fail(OXH_ERR_ARG, "oxh_batch_barostat_move: slot pools keep their boxes");
	if(!u4) return fail(OXH_ERR_ARG, "oxh_batch_barostat_move: null random numbers");
	oxh::Batch &b = B->b;
	if(b.h_sys[sys].has_ext)
		return fail(OXH_ERR_ARG, "oxh_batch_barostat_move: system %d has external forces, whose potential the move would need", sys);
	const int f = B->first[sys], n = B->n[sys];
	const oxh_sys S0 = b.h_sys[sys];
	const double T = S0.T;
	double U0 = 0, U1 = 0;
	TRY(system_energy(B, sys, U0));
	// the proposal: every side moved by delta_L (u - 1/2), the same u for
	// all three when isotropic
	const double L0[3] = {S0.box[0], S0.box[1], S0.box[2]};
	double L1[3];
	for(int k = 0; k < 3; k++) L1[k] = L0[k] + delta_L * (u4[isotropic ? 0 : k] - 0.5);
	std::vector<double4> p0;
	std::vector<int> o;
	TRY(download(B, p0, b.pos, f, n));
	TRY(download(B, o, b.orig, f, n));
	TRY(sync_and_check(B));
	std::vector<double4> p1 = p0;
	if(molecular) {
		// every strand moved with its centre of mass (the mean of its
		// particles' positions), which scales with the box
		const std::vector<int> &sid = B->strand_of[sys];
		const int ns = B->n_strands[sys];
		std::vector<double> com(3 * (size_t) ns, 0.0);
		std::vector<int> cnt(ns, 0);
		for(int k = 0; k < n; k++) {
			const int m = sid[o[k]];
			com[3 * m] += p0[k].x;
			com[3 * m + 1] += p0[k].y;
			com[3 * m + 2] += p0[k].z;
			cnt[m]++;
		}
		for(int k = 0; k < n; k++) {
			const int m = sid[o[k]];
			p1[k].x += com[3 * m] / cnt[m] * (L1[0] / L0[0] - 1.0);
			p1[k].y += com[3 * m + 1] / cnt[m] * (L1[1] / L0[1] - 1.0);
			p1[k].z += com[3 * m + 2] / cnt[m] * (L1[2] / L0[2] - 1.0);
		}
	}
	else
		for(int k = 0; k < n; k++) {
			p1[k].x *= L1[0] / L0[0];
			p1[k].y *= L1[1] / L0[1];
			p1[k].z *= L1[2] / L0[2];
		}
	if(b.graph_exec) {
		(void) hipGraphExecDestroy(b.graph_exec);
		b.graph_exec = nullptr;
	}
	auto apply = [&](const double *L, const std::vector<double4> &pos, const oxh_sys *restore) -> oxh_status {
		if(restore) {
			b.h_sys[sys] = *restore;
			HIPCHK(hipMemcpyAsync(b.d_sys + sys, restore, sizeof(oxh_sys), hipMemcpyHostToDevice, b.stream));
		}
		else TRY(set_system_box(B, sys, L));
		HIPCHK(hipMemcpyAsync(b.pos + f, pos.data(), (size_t) n * sizeof(double4), hipMemcpyHostToDevice, b.stream));
		oxh::launch_prepare_mirror(b.stream, b);
		TRY(sync_and_check(B));
		return oxh_batch_update_lists(B, 1);
	};
	oxh_status st = apply(L1, p1, nullptr);
	if(st == OXH_OK) st = system_energy(B, sys, U1);
	if(st != OXH_OK) {
		// a box too small for its cells, or a failed evaluation: back to the
		// state before the move
		const std::string err = g_err;
		TRY(apply(L0, p0, &S0));
		TRY(system_energy(B, sys, U0));
		g_err = err;
		return st;
	}
	const double V0 = L0[0] * L0[1] * L0[2], V1 = L1[0] * L1[1] * L1[2];
	const double n_obj = molecular ? (double) B->n_strands[sys] + 1.0 : (double) n;
	const double x = -((U1 - U0) + P * (V1 - V0) - n_obj * T * std::log(V1 / V0)) / T;
	const bool acc = std::exp(x) > u4[3];
	if(dE) *dE = U1 - U0;
	if(accepted) *accepted = acc ? 1 : 0;
	if(!acc) {
		TRY(apply(L0, p0, &S0));
		TRY(system_energy(B, sys, U0));
	}
	return OXH_OK;
}

oxh_status oxh_batch_box(oxh_batch *B, int sys, double *box3) {
	TRY(check_sys(B, sys));
	if(!box3) return fail(OXH_ERR_ARG, "null box buffer");
	for(int k = 0; k < 3; k++) box3[k] = B->b.h_sys[sys].box[k];
	return OXH_OK;
}

oxh_status oxh_batch_stats(oxh_batch *B, double *out6) {  // 12 doubles
	if(!B || !out6) return fail(OXH_ERR_ARG, "bad arguments");
	HIPCHK(hipSetDevice(B->b.device));
	std::vector<int2> nn;
	TRY(download(B, nn, B->b.nnbr, 0, B->b.N));
	TRY(sync_and_check(B));
	// mean over the systems' particles (a pool's free capacity holds no lists)
	double acc = 0;
	long long n_real = 0;
	for(const oxh_sys &S : B->b.h_sys) {
		for(int i = S.first; i < S.first + S.n; i++) acc += nn[i].x + nn[i].y;
		n_real += S.n;
	}
	out6[0] = B->b.N;
	out6[1] = B->b.max_neigh;
	out6[2] = n_real ? acc / n_real : 0.0;
	out6[3] = (double) B->device_bytes;
	out6[4] = B->b.lanes;
	out6[5] = B->b.force_block;
	out6[6] = B->b.far_lanes;
	out6[7] = B->b.far_block;
	out6[8] = B->persist ? 1.0 : (B->b.edge ? 2.0 : 0.0);
	out6[9] = B->b.nsys;
	out6[10] = B->b.pack ? 1.0 : 0.0;
	int far_back = 0;
	for(const oxh_sys &S : B->b.h_sys) far_back += S.far_back;
	out6[11] = far_back;
	return OXH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- checkpoints
// Layout: a header, then the device arrays in the batch's internal
// (possibly Hilbert-sorted) order, so that a run restored into a batch
// built from the same systems continues bitwise as if never interrupted.
namespace {
struct CheckpointHeader {
	char magic[8];            // "OXHCKPT2"
	int N, nsys, fp64, persist;
	long long step, since_sort, since_fix;
	int lanes, force_block, far_lanes, far_block, edge, edge_group, ecap, pad_;
	// FNV-1a of every system's parameters (range, box, cells, model
	// constants, integrator, thermostat and seed): a checkpoint restores
	// only into a batch built from the same systems
	unsigned long long sys_hash;
};

unsigned long long systems_fingerprint(const oxh_batch *B) {
	unsigned long long h = 1469598103934665603ull;
	for(const oxh_sys &S0 : B->b.h_sys) {
		oxh_sys S = S0;
		S.has_ext = 0;        // external forces are set again after a restore
		S.ext_step0 = 0;
		const unsigned char *c = (const unsigned char *) &S;
		for(size_t k = 0; k < sizeof(S); k++) {
			h ^= c[k];
			h *= 1099511628211ull;
		}
	}
	return h;
}

// every device array a checkpoint holds, with its byte size
std::vector<std::pair<void *, size_t>> ckpt_layout(oxh_batch *B) {
	oxh::Batch &b = B->b;
	const size_t N = b.N;
	std::vector<std::pair<void *, size_t>> a = {
			{b.pos, N * sizeof(double4)}, {b.vel, N * sizeof(double4)}, {b.Lm, N * sizeof(double4)}, {b.q64, N * sizeof(double4)},
			{b.rlist, N * sizeof(double4)}, {b.F, N * sizeof(float4)}, {b.T, N * sizeof(float4)}, {b.pstep, N * sizeof(long long)},
			{b.orig, N * sizeof(int)}, {b.bonds, N * sizeof(int2)}, {b.pinfo, N * sizeof(int2)},
			{b.img, N * sizeof(int4)}, {b.d_state, (size_t) b.nsys * sizeof(oxh_sys_state)}};
	if(b.fp64) {
		a.push_back({b.F64, N * sizeof(double4)});
		a.push_back({b.T64, N * sizeof(double4)});
	}
	if(b.blist) a.push_back({b.blist, N * sizeof(float4)});
	return a;
}
}  // namespace

extern "C" {

oxh_status oxh_batch_checkpoint_bytes(oxh_batch *B, size_t *bytes) {
	if(!B || !bytes) return fail(OXH_ERR_ARG, "bad arguments");
	if(B->pool) return fail(OXH_ERR_ARG, "checkpoints are taken of batches, not slot pools");
	size_t n = sizeof(CheckpointHeader);
	for(auto &x : ckpt_layout(B)) n += x.second;
	*bytes = n;
	return OXH_OK;
}

oxh_status oxh_batch_save(oxh_batch *B, void *buf, size_t bytes) {
	size_t need = 0;
	TRY(oxh_batch_checkpoint_bytes(B, &need));
	if(!buf || bytes < need) return fail(OXH_ERR_ARG, "checkpoint buffer of %zu bytes, %zu needed", bytes, need);
	oxh::Batch &b = B->b;
	HIPCHK(hipSetDevice(b.device));
	HIPCHK(hipStreamSynchronize(b.stream));
	CheckpointHeader h;
	std::memset(&h, 0, sizeof(h));
	std::memcpy(h.magic, "OXHCKPT2", 8);
	h.sys_hash = systems_fingerprint(B);
	h.N = b.N;
	h.nsys = b.nsys;
	h.fp64 = b.fp64;
	h.persist = B->persist;
	h.step = b.step;
	h.since_sort = B->since_sort;
	h.since_fix = B->since_fix;
	h.lanes = b.lanes;
	h.force_block = b.force_block;
	h.far_lanes = b.far_lanes;
	h.far_block = b.far_block;
	h.edge = b.edge;
	h.edge_group = b.edge_group;
	h.ecap = b.ecap;
	char *p = (char *) buf;
	std::memcpy(p, &h, sizeof(h));
	p += sizeof(h);
	for(auto &x : ckpt_layout(B)) {
		HIPCHK(hipMemcpyAsync(p, x.first, x.second, hipMemcpyDeviceToHost, b.stream));
		p += x.second;
	}
	return sync_and_check(B);
}

oxh_status oxh_batch_restore(oxh_batch *B, const void *buf, size_t bytes) {
	size_t need = 0;
	TRY(oxh_batch_checkpoint_bytes(B, &need));
	if(!buf || bytes < need) return fail(OXH_ERR_ARG, "checkpoint of %zu bytes, %zu expected", bytes, need);
	CheckpointHeader h;
	std::memcpy(&h, buf, sizeof(h));
	oxh::Batch &b = B->b;
	if(std::memcmp(h.magic, "OXHCKPT2", 8) != 0) return fail(OXH_ERR_ARG, "not a checkpoint of this library version");
	if(h.N != b.N || h.nsys != b.nsys || h.fp64 != (int) b.fp64 || h.persist != (int) B->persist)
		return fail(OXH_ERR_ARG, "checkpoint of another batch (N %d, %d systems) than this one (N %d, %d systems)", h.N, h.nsys, b.N,
				b.nsys);
	if(h.sys_hash != systems_fingerprint(B))
		return fail(OXH_ERR_ARG, "checkpoint of systems with other sizes, boxes or parameters than this batch's");
	HIPCHK(hipSetDevice(b.device));
	HIPCHK(hipStreamSynchronize(b.stream));
	// the checkpointed launch shape, so that forces round as they did
	if(h.edge && b.ecap == 0) TRY(alloc_edges(B, h.ecap > 0 ? h.ecap : std::min(b.max_neigh, 64)));
	if(!h.edge && b.ecap > 0) TRY(alloc_edges(B, 0));
	b.lanes = h.lanes;
	b.force_block = h.force_block;
	b.far_lanes = h.far_lanes;
	b.far_block = h.far_block;
	b.edge = h.edge;
	b.edge_group = h.edge_group;
	const char *p = (const char *) buf + sizeof(h);
	for(auto &x : ckpt_layout(B)) {
		HIPCHK(hipMemcpyAsync(x.first, p, x.second, hipMemcpyHostToDevice, b.stream));
		p += x.second;
	}
	HIPCHK(hipStreamSynchronize(b.stream));
	b.step = h.step;
	B->since_sort = h.since_sort;
	B->since_fix = h.since_fix;
	if(b.graph_exec) {
		(void) hipGraphExecDestroy(b.graph_exec);
		b.graph_exec = nullptr;
	}
	oxh::launch_prepare_mirror(b.stream, b);
	oxh::launch_slot_of(b.stream, b);
	// the lists of the checkpointed run, rebuilt exactly from their build
	// positions; the rebuild flags and counters are the checkpoint's
	std::vector<oxh_sys_state> saved(b.nsys);
	HIPCHK(hipMemcpyAsync(saved.data(), b.d_state, b.nsys * sizeof(oxh_sys_state), hipMemcpyDeviceToHost, b.stream));
	HIPCHK(hipStreamSynchronize(b.stream));
	TRY(rebuild_from_rlist(B));
	HIPCHK(hipMemcpyAsync(b.d_state, saved.data(), b.nsys * sizeof(oxh_sys_state), hipMemcpyHostToDevice, b.stream));
	return sync_and_check(B);
}

oxh_status oxh_batch_time_forces(oxh_batch *B, int n_launches, double *avg_ms) {
	if(!B || n_launches <= 0 || !avg_ms) return fail(OXH_ERR_ARG, "bad arguments");
	HIPCHK(hipSetDevice(B->b.device));
	hipEvent_t e0, e1;
	HIPCHK(hipEventCreate(&e0));
	HIPCHK(hipEventCreate(&e1));
	oxh::launch_forces(B->b.stream, B->b, false);   // warm
	HIPCHK(hipEventRecord(e0, B->b.stream));
	for(int k = 0; k < n_launches; k++) oxh::launch_forces(B->b.stream, B->b, false);
	HIPCHK(hipEventRecord(e1, B->b.stream));
	// leave the forces of the engine that runs the steps
	if(B->persist) enqueue_forces(B);
	HIPCHK(hipEventSynchronize(e1));
	float ms = 0.f;
	HIPCHK(hipEventElapsedTime(&ms, e0, e1));
	(void) hipEventDestroy(e0);
	(void) hipEventDestroy(e1);
	*avg_ms = ms / n_launches;
	return sync_and_check(B);
}

oxh_status oxh_hilbert_keys(int device, int n, const unsigned int *xyzb, unsigned int *keys) {
	if(n <= 0) return OXH_OK;
	HIPCHK(hipSetDevice(device));
	void *d_in = nullptr, *d_out = nullptr;
	HIPCHK(hipMalloc(&d_in, (size_t) n * 16));
	HIPCHK(hipMalloc(&d_out, (size_t) n * 4));
	HIPCHK(hipMemcpy(d_in, xyzb, (size_t) n * 16, hipMemcpyHostToDevice));
	oxh_internal_hilbert_probe(n, d_in, d_out, nullptr);
	HIPCHK(hipDeviceSynchronize());
	HIPCHK(hipMemcpy(keys, d_out, (size_t) n * 4, hipMemcpyDeviceToHost));
	(void) hipFree(d_in);
	(void) hipFree(d_out);
	return OXH_OK;
}

}  // extern "C"

__global__ void k_philox_probe(int n, const uint4 *ctr, const uint2 *key, uint4 *out) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i < n) out[i] = oxh::philox4x32_10(ctr[i], key[i]);
}

extern "C" oxh_status oxh_philox(int device, int n, const unsigned int *ctr4, const unsigned int *key2, unsigned int *out4) {
	if(n <= 0) return OXH_OK;
	HIPCHK(hipSetDevice(device));
	uint4 *dc = nullptr, *dout = nullptr;
	uint2 *dk = nullptr;
	HIPCHK(hipMalloc(&dc, (size_t) n * 16));
	HIPCHK(hipMalloc(&dk, (size_t) n * 8));
	HIPCHK(hipMalloc(&dout, (size_t) n * 16));
	HIPCHK(hipMemcpy(dc, ctr4, (size_t) n * 16, hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(dk, key2, (size_t) n * 8, hipMemcpyHostToDevice));
	k_philox_probe<<<(n + 255) / 256, 256>>>(n, dc, dk, dout);
	HIPCHK(hipDeviceSynchronize());
	HIPCHK(hipMemcpy(out4, dout, (size_t) n * 16, hipMemcpyDeviceToHost));
	(void) hipFree(dc);
	(void) hipFree(dk);
	(void) hipFree(dout);
	return OXH_OK;
}

// STREAM-style copy: each thread moves float4s, grid-stride
__global__ void __launch_bounds__(256) k_copy_probe(size_t n, const float4 *__restrict__ src, float4 *__restrict__ dst) {
	for(size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x) dst[i] = src[i];
}

extern "C" oxh_status oxh_copy_bench(int device, long long bytes, int reps, double *gbps) {
	if(bytes < 16 || reps <= 0 || !gbps) return fail(OXH_ERR_ARG, "oxh_copy_bench: bad arguments");
	HIPCHK(hipSetDevice(device));
	const size_t n = (size_t) bytes / 16;
	float4 *a = nullptr, *b = nullptr;
	HIPCHK(hipMalloc(&a, n * 16));
	if(hipMalloc(&b, n * 16) != hipSuccess) {
		(void) hipFree(a);
		return fail(OXH_ERR_ALLOC, "oxh_copy_bench: allocation failed");
	}
	hipEvent_t e0, e1;
	HIPCHK(hipEventCreate(&e0));
	HIPCHK(hipEventCreate(&e1));
	const int grid = 256 * 8 * 4;   // 4 workgroups of 256 per SIMD-pair over 256 CUs
	HIPCHK(hipMemset(a, 0, n * 16));
	k_copy_probe<<<grid, 256>>>(n, a, b);   // warm
	HIPCHK(hipEventRecord(e0, 0));
	for(int r = 0; r < reps; r++) k_copy_probe<<<grid, 256>>>(n, a, b);
	HIPCHK(hipEventRecord(e1, 0));
	HIPCHK(hipEventSynchronize(e1));
	float ms = 0.f;
	HIPCHK(hipEventElapsedTime(&ms, e0, e1));
	*gbps = 2.0 * (double) n * 16 * reps / (ms * 1e-3) / 1e9;
	(void) hipEventDestroy(e0);
	(void) hipEventDestroy(e1);
	(void) hipFree(a);
	(void) hipFree(b);
	return OXH_OK;
}

// ------------------------------------------------- forces without allocation
// oxh_interaction_compute_forces_noalloc: one system's pair forces with every
// buffer beyond the caller's inputs and outputs in `scratch`, a fixed device
// array of this library (one copy per GPU, held until the process ends).
// Its regions, each 256-byte aligned: the system's parameters and state
// word, the rebuild flag, the topology staging, the FP64 state and list
// positions, the FP32 mirror, F / T in both precisions, bonds, particle
// info, original indices and the Verlet list (OXH_NOALLOC_MAX_NEIGH
// entries per particle).  Host code takes its address per device from
// hipGetSymbolAddress and never uses the array's name as a device address.
namespace {

constexpr int NA_N = OXH_NOALLOC_MAX_N;
constexpr int NA_NEIGH = OXH_NOALLOC_MAX_NEIGH;

constexpr size_t na_up(size_t x) { return (x + 255) & ~(size_t) 255; }

struct NaLayout {
	static constexpr size_t sys = 0;
	static constexpr size_t state = na_up(sys + sizeof(oxh_sys));
	static constexpr size_t any = na_up(state + sizeof(oxh_sys_state));
	static constexpr size_t topo = na_up(any + sizeof(int));
	static constexpr size_t pos = na_up(topo + NA_N * sizeof(int4));
	static constexpr size_t q64 = pos + NA_N * sizeof(double4);
	static constexpr size_t rlist = q64 + NA_N * sizeof(double4);
	static constexpr size_t back64 = rlist + NA_N * sizeof(double4);
	static constexpr size_t F64 = back64 + NA_N * sizeof(double4);
	static constexpr size_t T64 = F64 + NA_N * sizeof(double4);
	static constexpr size_t q32 = T64 + NA_N * sizeof(double4);
	static constexpr size_t bpos = q32 + NA_N * sizeof(float4);
	static constexpr size_t F = bpos + NA_N * sizeof(float4);
	static constexpr size_t T = F + NA_N * sizeof(float4);
	static constexpr size_t bonds = T + NA_N * sizeof(float4);
	static constexpr size_t pinfo = bonds + NA_N * sizeof(int2);
	static constexpr size_t nnbr = pinfo + NA_N * sizeof(int2);
	static constexpr size_t orig = nnbr + NA_N * sizeof(int2);
	static constexpr size_t nbr = na_up(orig + NA_N * sizeof(int));
	static constexpr size_t end = na_up(nbr + (size_t) NA_N * NA_NEIGH * sizeof(int));
};
static_assert(NaLayout::end == OXH_NOALLOC_SCRATCH_BYTES, "scratch size in include/oxh.h");

static __device__ __align__(256) unsigned char scratch[NaLayout::end];

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mutex;
unsigned char *g_na_ptr[NA_MAX_DEVICES] = {};
// host staging of the topology (n3, n5, btype) and of the state word read back
int4 g_na_topo[NA_N];
oxh_sys_state g_na_state;

// the device address of `scratch` on `device` (current), looked up once
oxh_status na_scratch(int device, unsigned char **out) {
	if(!g_na_ptr[device]) {
		void *p = nullptr;
		hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if(e != hipSuccess || !p) return fail(OXH_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(e));
		g_na_ptr[device] = (unsigned char *) p;
	}
	*out = g_na_ptr[device];
	return OXH_OK;
}

// parameters, state word and rebuild flag; the caller's state and the
// staged topology into the scratch arrays, with the FP32 mirror
__global__ void k_na_load(int N, const oxh_sys S, const int4 *__restrict__ topo, const double4 *__restrict__ upos,
		const double4 *__restrict__ uq, oxh_sys *__restrict__ d_sys, oxh_sys_state *__restrict__ st, int *__restrict__ any,
		double4 *__restrict__ pos, double4 *__restrict__ q64, float4 *__restrict__ q32, float4 *__restrict__ bpos,
		double4 *__restrict__ back64, int2 *__restrict__ bonds, int2 *__restrict__ pinfo, int *__restrict__ orig) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i == 0) {
		*d_sys = S;
		st->K_t = st->K_r = 0.0;
		st->need_rebuild = 1;
		st->error = 0;
		st->n_rebuilds = 0;
		*any = 1;
	}
	if(i >= N) return;
	const int4 t = topo[i];
	const int2 bd = make_int2(t.x, t.y);
	bonds[i] = bd;
	pinfo[i] = make_int2(t.z, 0);
	orig[i] = i;
	double4 r = upos[i];
	r.w = 0.0;
	const double4 q = uq[i];
	pos[i] = r;
	q64[i] = q;
	oxh::write_mirror(i, r, q, S, bd, q32, bpos, back64);
}

}  // namespace

extern "C" oxh_status oxh_interaction_compute_forces_noalloc(int device, const oxh_system_desc *desc, const double *d_pos,
		const double *d_quat, void *d_forces, void *d_torques, void *stream) {
	if(!desc || !d_pos || !d_quat || !d_forces || !d_torques)
		return fail(OXH_ERR_ARG, "oxh_interaction_compute_forces_noalloc: bad arguments");
	if(device < 0 || device >= NA_MAX_DEVICES) return fail(OXH_ERR_ARG, "oxh_interaction_compute_forces_noalloc: bad device %d", device);
	const int N = desc->N;
	if(N > NA_N)
		return fail(OXH_ERR_ARG, "oxh_interaction_compute_forces_noalloc: %d nucleotides, at most %d (OXH_NOALLOC_MAX_N)", N, NA_N);
	if(desc->backend_precision != 0 && desc->backend_precision != 1)
		return fail(OXH_ERR_ARG, "oxh_interaction_compute_forces_noalloc: backend_precision must be 0 or 1");
	oxh_sys S;
	double rv = 0;
	TRY(sys_setup(nullptr, *desc, 0, S, rv, false));
	S.first = 0;
	S.n = N;
	S.use_cells = 0;   // all-pairs list build
	S.cell_first = S.ncell = S.stencil = 0;
	S.has_ext = 0;
	std::lock_guard<std::mutex> lock(g_na_mutex);
	HIPCHK(hipSetDevice(device));
	unsigned char *base = nullptr;
	TRY(na_scratch(device, &base));
	hipStream_t st = (hipStream_t) stream;
	for(int i = 0; i < N; i++) g_na_topo[i] = make_int4(desc->n3[i], desc->n5[i], desc->btype[i], 0);
	HIPCHK(hipMemcpyAsync(base + NaLayout::topo, g_na_topo, (size_t) N * sizeof(int4), hipMemcpyHostToDevice, st));

	oxh::Batch b;
	b.device = device;
	b.stream = st;
	b.N = N;
	b.nsys = 1;
	b.d_sys = (oxh_sys *) (base + NaLayout::sys);
	b.d_state = (oxh_sys_state *) (base + NaLayout::state);
	b.any_rebuild = (int *) (base + NaLayout::any);
	b.pos = (double4 *) (base + NaLayout::pos);
	b.q64 = (double4 *) (base + NaLayout::q64);
	b.rlist = (double4 *) (base + NaLayout::rlist);
	b.back64 = (double4 *) (base + NaLayout::back64);
	b.F64 = (double4 *) (base + NaLayout::F64);
	b.T64 = (double4 *) (base + NaLayout::T64);
	b.q32 = (float4 *) (base + NaLayout::q32);
	b.bpos = (float4 *) (base + NaLayout::bpos);
	b.F = (float4 *) (base + NaLayout::F);
	b.T = (float4 *) (base + NaLayout::T);
	b.bonds = (int2 *) (base + NaLayout::bonds);
	b.pinfo = (int2 *) (base + NaLayout::pinfo);
	b.nnbr = (int2 *) (base + NaLayout::nnbr);
	b.orig = (int *) (base + NaLayout::orig);
	b.nbr = (int *) (base + NaLayout::nbr);
	b.nstride = N;
	b.max_neigh = NA_NEIGH;
	b.fp64 = desc->backend_precision == 1;
	b.model = desc->interaction == OXH_INTERACTION_DNA2 ? 0 : 1;
	b.lanes = 8;
	b.far_lanes = 0;
	b.force_block = 256;
	b.pack = false;

	const int bs = 256, g = (N + bs - 1) / bs;
	k_na_load<<<g, bs, 0, st>>>(N, S, (const int4 *) (base + NaLayout::topo), (const double4 *) d_pos, (const double4 *) d_quat,
			b.d_sys, b.d_state, b.any_rebuild, b.pos, b.q64, b.q32, b.bpos, b.back64, b.bonds, b.pinfo, b.orig);
	oxh::launch_build_list(st, b, 1, b.pos);
	oxh::launch_forces(st, b, false);
	HIPCHK(hipMemcpyAsync(&g_na_state, b.d_state, sizeof(oxh_sys_state), hipMemcpyDeviceToHost, st));
	HIPCHK(hipStreamSynchronize(st));
	HIPCHK(hipGetLastError());
	if(g_na_state.error & oxh::ERR_LIST)
		return fail(OXH_ERR_SYSTEM, "oxh_interaction_compute_forces_noalloc: a particle has more than %d neighbours (OXH_NOALLOC_MAX_NEIGH)",
				NA_NEIGH);
	if(b.fp64)
		k_add_to_caller<double4><<<g, bs, 0, st>>>(N, b.F64, b.T64, b.orig, b.pinfo, b.d_sys, (double4 *) d_forces, (double4 *) d_torques);
	else
		k_add_to_caller<float4><<<g, bs, 0, st>>>(N, b.F, b.T, b.orig, b.pinfo, b.d_sys, (float4 *) d_forces, (float4 *) d_torques);
	HIPCHK(hipGetLastError());
	return OXH_OK;
}
