// Internal interface between the C ABI (oxh_api.hip) and the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "oxh_types.h"

namespace oxh {

// A batch: one or more independent systems packed into one set of SoA
// device arrays (system s owns particles [first, first + n)).  The plain
// single-system path is a batch of one.
struct Batch {
	int device = 0;
	hipStream_t stream = nullptr;
	int N = 0;          // particles in the batch
	int nsys = 0;

	// per-particle state (FP64 integration state, FP32 kernel mirror)
	double4 *pos = nullptr, *vel = nullptr, *Lm = nullptr, *q64 = nullptr, *rlist = nullptr;
	float4 *q32 = nullptr, *F = nullptr, *T = nullptr;
	// all-FP64 mode (backend_precision = double): forces and torques in FP64
	bool fp64 = false;
	double4 *F64 = nullptr, *T64 = nullptr;
	double4 *back64 = nullptr;   // FP64 backbone offsets (POS_MM_BACK1 a1 + POS_MM_BACK2 a2), written with the mirror
	float4 *bpos = nullptr;   // backbone-site position wrapped into the box (FP32), .w = Debye-Hueckel charge factor
	int2 *bonds = nullptr;   // (n3, n5) batch indices, -1 = none
	int2 *pinfo = nullptr;   // (btype, system)
	int *orig = nullptr;     // index of the particle inside its system as uploaded
	int *slot_of = nullptr;  // inverse of orig: slot_of[first + orig[i]] = i
	int4 *img = nullptr;     // periodic image counts (absolute position = pos + img * box)
	long long *pstep = nullptr;  // per-particle step counter (RNG counter)

	// neighbour list, row major: particle i owns nbr[i * max_neigh + k];
	// close pairs (may need every term) fill it from the front, far pairs
	// (Debye-Hueckel only) from the back; nnbr[i] = (n_close, n_far)
	int *nbr = nullptr;
	int2 *nnbr = nullptr;
	int nstride = 0, max_neigh = 0;
	int lanes = 8;             // threads per nucleotide in the force kernel
	int far_lanes = 0;         // threads per nucleotide of the separate far-pair kernel (0 = far pairs inline)
	int far_block = 256;

	// edge (half-list) form of the close pairs, rebuilt with the lists when
	// ecap > 0 and used by the step's force evaluation when edge is set.
	// Row i (width ecap) of eown holds i's close neighbours j > i in list
	// order: the pair (i, eown[i][m]) is evaluated once, into slot
	// i * ecap + m of E1 (force on j, half energy), E2 (lab torque on i, half
	// HB energy), E3 (lab torque on j).  Row i of ein holds the slots of i's
	// pairs with j < i.  ecount[i] = (owned, incoming).  B1..B3[i]: the bond
	// i -> n3 likewise.  edge_group: particles per wave of the pair kernel.
	bool edge = false;
	int ecap = 0, edge_group = 16;
	int *eown = nullptr, *ein = nullptr;
	int2 *ecount = nullptr;
	float4 *E1 = nullptr, *E2 = nullptr, *E3 = nullptr, *B1 = nullptr, *B2 = nullptr, *B3 = nullptr;

	// cells (systems with use_cells), fixed capacity per cell
	int ncell_total = 0, cell_cap = 0;
	bool cell_build = false;   // cell-centric list build (k_build_cells)
	// (cell_raw: slots in binning order; cell_items: the same sorted by
	// particle index; cell_of: a binned particle's cell)
	int *cell_count = nullptr, *cell_items = nullptr, *cell_owner = nullptr, *cell_raw = nullptr, *cell_of = nullptr;
	float4 *pos32 = nullptr;   // wrapped FP32 centres, written by the cell fill of a rebuild

	// per-system parameters and device-written state
	oxh_sys *d_sys = nullptr;
	oxh_sys_state *d_state = nullptr;
	std::vector<oxh_sys> h_sys;
	// slot pools: steps each system may take in the current chunk (step k of
	// the chunk runs for system s iff k < limit[s]); nullptr = no limit.
	// Particles [first + n, first + capacity) of a slot are padding, which
	// every kernel skips.
	int *limit = nullptr;
	// set when some system's skin check (or the host) asks for a list build
	// in this step, cleared by the step's second half; the per-step cell and
	// list kernels return at once while it is clear
	int *any_rebuild = nullptr;

	// reductions: blocks of at most RED_CHUNK particles, each inside one system
	int nred = 0;
	int3 *red_blocks = nullptr;        // (system, begin, end)
	int2 *red_sys_blocks = nullptr;    // per system: (first block, n blocks)
	double *red_partial = nullptr;     // nred * RED_W
	double *red_out = nullptr;         // nsys * RED_W
	int *red_done = nullptr;           // per system: blocks finished (Bussi's last-block reduction)

	// optional outputs
	double *esplit = nullptr;   // 8 * N
	double *stress = nullptr;   // 6 * N: each particle's half of its pairs' virial (xx yy zz xy xz yz)
	int *hbcount = nullptr;    // N
	bool want_hbcount = false;

	// external forces: CSR over (system first + original index), N + 1
	// offsets, records in the order the caller gave them per particle
	// backbone sites at the last list build (systems with far_back), the
	// bound of their moves before a rebuild
	float4 *blist = nullptr;
	int n_ext = 0;
	int *ext_off = nullptr;
	oxh_ext_rec *ext_rec = nullptr;
	// centre-of-mass groups of "com" forces: ext_grp[g] = (first, count)
	// of the group's com list and (first, count) of its reference list in
	// ext_idx (indices in the system, upload order); ext_com[g] holds the
	// two centres (6 doubles), written each force evaluation
	int n_grp = 0;
	int4 *ext_grp = nullptr;
	int *ext_grp_sys = nullptr;
	int *ext_idx = nullptr;
	double *ext_com = nullptr;

	// Hilbert sort scratch
	unsigned int *keys = nullptr, *keys_alt = nullptr;
	int *perm = nullptr, *perm_alt = nullptr, *inv_perm = nullptr;
	unsigned int *radix_hist = nullptr;
	void *scratch = nullptr;
	size_t scratch_bytes = 0;

	int force_block = 256;
	// force kernels: workgroups mapped to XCDs in contiguous ranges
	// (xcd_block; OXH_XCD=0 takes the hardware order)
	int xcd = 1;
	// workgroup engine with each system resident in LDS: the largest
	// system's size (0 = the engine reads and writes HBM every step)
	int persist_lds_n = 0;
	bool pack = false;   // close pairs' HB/CRST/CXST through LDS work lists (k_forces PACK)
	int model = 0;       // pair terms: 0 oxDNA2, 1 oxRNA (k_forces_rna; Debye-Hueckel per system)
	int any_cells = 0;
	int thermostat_mask = 0;     // bit t set if some system uses thermostat t
	int min_newtonian = 1;

	hipGraphExec_t graph_exec = nullptr;
	int graph_steps = 0;
	long long step = 0;          // steps run by this batch (host count)
	std::string last_error;
};

// oxh_sys_state.error bits
constexpr int ERR_BOND = 1, ERR_LIST = 2, ERR_CELLS = 4, ERR_EDGE = 8;
constexpr int ERR_EXT = 16;   // a particle where an external force is undefined (the reference throws)
constexpr int ERR_OVERFLOW = ERR_LIST | ERR_CELLS | ERR_EDGE;   // fixed by growing a buffer and replaying

constexpr int RED_CHUNK = 2048;
constexpr int RED_W = 8;   // reduction width: U, Uhb, Kt, Kr, vx, vy, vz, count

// workgroup-resident engine: one workgroup of `block` threads (256 or
// 512) per system, for batches whose systems all have at most
// PERSIST_MAX_N particles and no cells
constexpr int PERSIST_MAX_N = 512;

// lanes per nucleotide of the workgroup-resident force phase: the largest
// power of two <= 16 with n * K <= block
__host__ __device__ inline int persist_lanes(int n, int block) {
	int k = 16;
	while(k > 1 && k * n > block) k >>= 1;
	return k;
}

void launch_persist(hipStream_t st, const Batch &b, int nsteps, int block);
// dynamic LDS of the LDS-resident workgroup engine for systems of at most
// nmax nucleotides
size_t persist_lds_bytes(int nmax);
// largest system the LDS-resident engine takes: its dynamic LDS plus the
// engine's static LDS within 64 KiB per workgroup
constexpr int PERSIST_LDS_MAX_N = 80;
void launch_persist_forces(hipStream_t st, const Batch &b, bool split, int block);

// k = index of the step inside the current chunk (checked against limit);
// k < 0: forces of every system's current state, whatever its limit
void launch_forces(hipStream_t st, const Batch &b, bool split, int k = -1);
// list builds read positions from src (the current positions, or rlist to
// rebuild exactly the lists of the last build)
void launch_cells(hipStream_t st, const Batch &b, const double4 *src);
void launch_cells_fill_sort(hipStream_t st, const Batch &b);
// (and the edge rows of rebuilt systems, when b.ecap > 0)
void launch_build_list(hipStream_t st, const Batch &b, int force_all, const double4 *src);
void launch_first_step(hipStream_t st, const Batch &b, int k);
void launch_second_step(hipStream_t st, const Batch &b, int k);
// second half of step k + first half of step k + 1
void launch_second_first(hipStream_t st, const Batch &b, int k);
void launch_bussi(hipStream_t st, const Batch &b, int k);
void launch_reduce(hipStream_t st, const Batch &b);
// velocities ~ N(0, T) for particles [i0, i0 + n)
void launch_refresh_vel(hipStream_t st, const Batch &b, int i0, int n);
void launch_prepare_mirror(hipStream_t st, const Batch &b);
int launch_hilbert_sort(hipStream_t st, Batch &b);
// adds every particle's external forces to F (F64) of the current state
void launch_ext_forces(hipStream_t st, const Batch &b);
// slot_of from orig (after an upload or a sort)
void launch_slot_of(hipStream_t st, const Batch &b);
// oxRNA / oxRNA2 pair forces (written, not accumulated) into F / T, or
// F64 / T64 in the all-FP64 form; b.lanes threads per nucleotide
void launch_rna_forces(hipStream_t st, const Batch &b, bool split, int k);
// the FFS per-pair quantities of oxRNA systems (OXH_PAIR_HB_ENERGY,
// _BASE_DISTANCE, _CENTRE_DISTANCE)
void launch_rna_pair_quantities(hipStream_t st, const Batch &b, int sys, int kind, int n, const int4 *d_pairs, double *d_out);
// per-pair FFS order-parameter quantities (OXH_PAIR_*) of system sys:
// pairs (p, q, same strand, unused) in the system's upload order
void launch_pair_quantities(hipStream_t st, const Batch &b, int sys, int kind, int n, const int4 *d_pairs, double *d_out);

}  // namespace oxh
