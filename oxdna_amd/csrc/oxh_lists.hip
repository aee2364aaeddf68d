// Neighbour lists and Hilbert ordering on gfx950.
//
// Verlet list: every j != i of the same system, not bonded to i, with
// |r_ij|^2 < r_verlet^2, r_verlet = r_cut + 2 skin; rebuilt when a particle
// moved more than the skin since the last build (reference:
// src/CUDA/Lists/CUDA_simple_verlet.cuh:14-95,
// src/CUDA/Lists/CUDASimpleVerletList.cu:96-289).  The predicate is
// evaluated in FP64 with a fixed operation sequence (explicit fma) so the
// pair set is reproducible bit for bit on any IEEE FP64 machine.  Cells
// hold a fixed number of slots and are sorted after filling, so list
// contents and order are deterministic (the reference fills cells with
// atomicInc slots in arbitrary order).
//
// Hilbert ordering (reference: src/CUDA/CUDA_sort.cu, with thrust's
// sort_by_key): key = Hilbert index of the particle's cell after
// J. Skilling, "Programming the Hilbert curve" (AIP Conf. Proc. 707, 2004),
// sorted by a hand-written stable LSD radix sort with wave64 ballots.
#include <algorithm>
#include <type_traits>
#include <utility>

#include "oxh_device.h"
#include "oxh_internal.h"
#include "oxh_list.h"

namespace oxh {

__global__ void k_cells_clear(int ncell, const int *__restrict__ owner, const oxh_sys_state *__restrict__ st, int *__restrict__ count) {
	int c = blockIdx.x * blockDim.x + threadIdx.x;
	if(c >= ncell) return;
	if(st[owner[c]].need_rebuild) count[c] = 0;
}

// Bins flagged systems' particles: cell slots in arrival order (raw), the
// particle's cell in cell_of, its wrapped FP32 centre in pos32.
__global__ void k_cells_fill(int N, const double4 *__restrict__ pos, const int2 *__restrict__ pinfo, const oxh_sys *__restrict__ systems,
		oxh_sys_state *__restrict__ st, int *__restrict__ count, int *__restrict__ raw, int *__restrict__ cell_of, int cap,
		float4 *__restrict__ pos32, const int *__restrict__ any, const float4 *__restrict__ bpos, float4 *__restrict__ blist) {
	if(!*any) return;
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= N) return;
	int s = pinfo[i].y;
	if(!st[s].need_rebuild) return;
	const oxh_sys &S = systems[s];
	if(!S.use_cells || padding(i, S)) return;
	double4 p = pos[i];
	int cx = cell_coord(p.x, S.inv_box[0], S.nc[0]);
	int cy = cell_coord(p.y, S.inv_box[1], S.nc[1]);
	int cz = cell_coord(p.z, S.inv_box[2], S.nc[2]);
	int c = S.cell_first + (cz * S.nc[1] + cy) * S.nc[0] + cx;
	pos32[i] = make_float4((float) (p.x - S.box[0] * floor(p.x * S.inv_box[0])), (float) (p.y - S.box[1] * floor(p.y * S.inv_box[1])),
			(float) (p.z - S.box[2] * floor(p.z * S.inv_box[2])), 0.f);
	cell_of[i] = c;
	// a build from the current state records the backbone sites it
	// filters far pairs with (a rebuild from rlist keeps those of its build)
	if(blist && S.far_back) blist[i] = bpos[i];
	int slot = atomicAdd(&count[c], 1);
	if(slot < cap) raw[(size_t) c * cap + slot] = i;
	else atomicOr(&st[s].error, ERR_CELLS);
}

// Puts every binned particle at its rank in its cell: items[c][rank] with
// rank = number of particles of the cell with a smaller index, so a cell's
// items are sorted by particle index whatever order the atomics gave.
__global__ void k_cells_rank(int N, const int2 *__restrict__ pinfo, const oxh_sys *__restrict__ systems,
		const oxh_sys_state *__restrict__ st, const int *__restrict__ count, const int *__restrict__ raw, const int *__restrict__ cell_of,
		int *__restrict__ items, int cap, const int *__restrict__ any) {
	if(!*any) return;
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= N) return;
	int s = pinfo[i].y;
	if(!st[s].need_rebuild) return;
	const oxh_sys &S = systems[s];
	if(!S.use_cells || padding(i, S)) return;
	const int c = cell_of[i];
	const int n = min(count[c], cap);
	const int *r = raw + (size_t) c * cap;
	int rank = 0;
	bool stored = false;
	for(int m = 0; m < n; m++) {
		const int v = r[m];
		rank += v < i;
		stored |= v == i;
	}
	if(stored) items[(size_t) c * cap + rank] = i;
}

// fill + rank of flagged systems' cells; the counters must be zero
void launch_cells_fill_sort(hipStream_t st, const Batch &b) {
	if(!b.any_cells) return;
	int bs = 256;
	k_cells_fill<<<(b.N + bs - 1) / bs, bs, 0, st>>>(b.N, b.pos, b.pinfo, b.d_sys, b.d_state, b.cell_count, b.cell_raw, b.cell_of,
			b.cell_cap, b.pos32, b.any_rebuild, b.bpos, b.blist);
	k_cells_rank<<<(b.N + bs - 1) / bs, bs, 0, st>>>(b.N, b.pinfo, b.d_sys, b.d_state, b.cell_count, b.cell_raw, b.cell_of, b.cell_items,
			b.cell_cap, b.any_rebuild);
}

void launch_cells(hipStream_t st, const Batch &b, const double4 *src) {
	if(!b.any_cells) return;
	int bs = 256;
	k_cells_clear<<<(b.ncell_total + bs - 1) / bs, bs, 0, st>>>(b.ncell_total, b.cell_owner, b.d_state, b.cell_count);
	k_cells_fill<<<(b.N + bs - 1) / bs, bs, 0, st>>>(b.N, src, b.pinfo, b.d_sys, b.d_state, b.cell_count, b.cell_raw, b.cell_of,
			b.cell_cap, b.pos32, b.any_rebuild, b.bpos, src == b.pos ? b.blist : nullptr);
	k_cells_rank<<<(b.N + bs - 1) / bs, bs, 0, st>>>(b.N, b.pinfo, b.d_sys, b.d_state, b.cell_count, b.cell_raw, b.cell_of, b.cell_items,
			b.cell_cap, b.any_rebuild);
}

__global__ void __launch_bounds__(256) k_build_list(int N, const double4 *__restrict__ pos, const int2 *__restrict__ bonds,
		const int2 *__restrict__ pinfo, const oxh_sys *__restrict__ systems, oxh_sys_state *__restrict__ st,
		const int *__restrict__ count, const int *__restrict__ items, int cap, int *__restrict__ nbr, int2 *__restrict__ nnbr,
		int max_neigh, double4 *__restrict__ rlist, int force_all, const float4 *__restrict__ pos32, const int *__restrict__ any,
		const float4 *__restrict__ blist) {
	constexpr int K = BUILD_LANES;
	if(!force_all && !*any) return;
	// grid-stride over the K-lane groups: a step without a rebuild only
	// dispatches a bounded number of workgroups; the stride is a multiple of
	// the wave, so every group stays whole and aligned
	const size_t stride = (size_t) gridDim.x * blockDim.x;
	for(size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x; t < (size_t) N * K; t += stride) {
		const int i = (int) (t / K);
		const int l = (int) (t % K);
		const int s = pinfo[i].y;
		if(!force_all && !st[s].need_rebuild) continue;
		if(padding(i, systems[s])) continue;
		build_row<K>(i, l, s, pos, bonds, systems, st, count, items, cap, nbr, nnbr, max_neigh, rlist, pos32, blist);
	}
}

// Edge rows of particle i, from its finished list row (the half-list form
// of the close pairs; the reference's edge lists, src/CUDA/Lists/
// CUDA_simple_verlet.cuh:97-175, hold each unordered pair once as well).
// Owned pairs (j > i) are compacted in row order; for each pair with j < i
// the slot that j writes is i's rank among j's owned pairs, i.e. the number
// of j's close neighbours above j that come before i in j's row (rows are
// symmetric: the pair predicate gives the same result from both ends).
// Ballot compaction in row order, so the rows do not depend on timing.
// Rows longer than ecap, or a pair missing from its partner's row, set
// ERR_EDGE and leave the particle's edge rows empty.
__global__ void __launch_bounds__(256) k_edge_rows(int N, const int2 *__restrict__ pinfo, const oxh_sys *__restrict__ systems,
		oxh_sys_state *__restrict__ st, const int *__restrict__ nbr, const int2 *__restrict__ nnbr, int max_neigh, int ecap,
		int *__restrict__ eown, int *__restrict__ ein, int2 *__restrict__ ecount, int force_all, const int *__restrict__ any) {
	constexpr int K = BUILD_LANES;
	if(!force_all && !*any) return;
	const size_t stride = (size_t) gridDim.x * blockDim.x;
	const int g0 = (threadIdx.x & 63) & ~(K - 1);
	const uint64_t gmask = ((1ull << K) - 1ull) << g0;
	const uint64_t below = ((1ull << (threadIdx.x & 63)) - 1ull) & gmask;
	for(size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x; t < (size_t) N * K; t += stride) {
		const int i = (int) (t / K);
		const int l = (int) (t % K);
		const int s = pinfo[i].y;
		if(!force_all && !st[s].need_rebuild) continue;
		if(padding(i, systems[s])) continue;
		const int nc = nnbr[i].x;
		const int *row = nbr + (size_t) i * max_neigh;
		int no = 0, ni = 0;
		bool bad = false;
		for(int m0 = 0; m0 < nc; m0 += K) {
			const int m = m0 + l;
			const int j = m < nc ? row[m] : -1;
			const bool own = j > i, inc = j >= 0 && j < i;
			int slot = -1;
			if(inc) {
				const int *rj = nbr + (size_t) j * max_neigh;
				const int ncj = nnbr[j].x;
				int rank = 0, q = 0;
				for(; q < ncj; q++) {
					const int x = rj[q];
					if(x == i) break;
					rank += x > j;
				}
				if(q < ncj && rank < ecap) slot = j * ecap + rank;
				else bad = true;
			}
			const uint64_t bo = __ballot(own) & gmask, bi = __ballot(inc) & gmask;
			if(own) {
				const int k = no + __popcll(bo & below);
				if(k < ecap) eown[(size_t) i * ecap + k] = j;
			}
			if(inc) {
				const int k = ni + __popcll(bi & below);
				if(k < ecap) ein[(size_t) i * ecap + k] = slot;
			}
			no += __popcll(bo);
			ni += __popcll(bi);
		}
		bad = __ballot(bad) & gmask;
		if(l != 0) continue;
		if(bad || no > ecap || ni > ecap) {
			atomicOr(&st[s].error, ERR_EDGE);
			no = ni = 0;
		}
		ecount[i] = make_int2(no, ni);
	}
}

// Cell-centric list build (b.cell_build: every system with cells, not a
// pool, several particles per cell): one workgroup per home cell
// (grid-stride over the cells) stages the candidates of the stencil cells
// (k order, each cell's items in slot order, cells out of range of the whole
// home cell culled) in LDS once, and each wave then builds one home
// particle's row from them, 64 candidates per ballot.  Culled cells hold no
// pair within r_verlet, so the rows are those of build_row entry for entry,
// in the same order.  Home cells with more than CELL_CMAX candidates fall
// back to build_row.
constexpr int CELL_CMAX = 2048;

// BACK: some system keeps far pairs by backbone sites; for such systems
// the candidates' recorded backbone sites are staged instead of their
// centres, and a candidate's centre is read only when its site passes.
template<bool BACK>
__global__ void __launch_bounds__(256) k_build_cells(int ncell_total, const double4 *__restrict__ pos, const int2 *__restrict__ bonds,
		const oxh_sys *__restrict__ systems, oxh_sys_state *__restrict__ st, const int *__restrict__ cell_owner,
		const int *__restrict__ count, const int *__restrict__ items, int cap, int *__restrict__ nbr, int2 *__restrict__ nnbr,
		int max_neigh, double4 *__restrict__ rlist, int force_all, const float4 *__restrict__ pos32, const int *__restrict__ any,
		const float4 *__restrict__ blist) {
	// candidates: wrapped FP32 centre (or backbone site) with the
	// particle index in w
	__shared__ float4 cpos[CELL_CMAX];
	__shared__ int ccell[128], cbase[129];
	__shared__ int nkept;
	if(!force_all && !*any) return;
	const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
	const uint64_t below = (1ull << lane) - 1ull;
	for(int c = blockIdx.x; c < ncell_total; c += gridDim.x) {
		// every test below reads values that are the same for the whole
		// workgroup, so its threads take the same barriers
		const int s = cell_owner[c];
		if(!force_all && !st[s].need_rebuild) continue;
		const int nh = min(count[c], cap);
		if(nh == 0) continue;
		const oxh_sys &S = systems[s];
		if(tid < 64) {
			// lane l takes stencil cells 2l and 2l + 1 (k order); culled and
			// missing cells count 0 candidates; one wave scan numbers them
			const int lc = c - S.cell_first;
			const int cx = lc % S.nc[0], cy = (lc / S.nc[0]) % S.nc[1], cz = lc / (S.nc[0] * S.nc[1]);
			const int R = S.stencil, W = 2 * R + 1, NS = W * W * W;
			const float cull = (sqrtf(S.rv2f) + 1e-3f) * (sqrtf(S.rv2f) + 1e-3f);
			auto gap = [](int o, float wd) { return o > 0 ? (o - 1) * wd : (o < 0 ? (-o - 1) * wd : 0.f); };
			int cc[2] = {0, 0}, cnt[2] = {0, 0};
			for(int q = 0; q < 2; q++) {
				const int k = 2 * lane + q;
				if(k >= NS) continue;
				const int oz = k / (W * W) - R, oy = (k / W) % W - R, ox = k % W - R;
				const float gx = gap(ox, S.cell_w[0]), gy = gap(oy, S.cell_w[1]), gz = gap(oz, S.cell_w[2]);
				if(gx * gx + gy * gy + gz * gz > cull) continue;
				int z = cz + oz, y = cy + oy, x = cx + ox;
				z += (z < 0) ? S.nc[2] : (z >= S.nc[2] ? -S.nc[2] : 0);
				y += (y < 0) ? S.nc[1] : (y >= S.nc[1] ? -S.nc[1] : 0);
				x += (x < 0) ? S.nc[0] : (x >= S.nc[0] ? -S.nc[0] : 0);
				cc[q] = S.cell_first + (z * S.nc[1] + y) * S.nc[0] + x;
				cnt[q] = min(count[cc[q]], cap);
			}
			const int v = cnt[0] + cnt[1];
			int incl = v;
			for(int off = 1; off < 64; off <<= 1) {
				const int u = __shfl_up(incl, off);
				if(lane >= off) incl += u;
			}
			const int excl = incl - v;
			ccell[2 * lane] = cc[0];
			ccell[2 * lane + 1] = cc[1];
			cbase[2 * lane] = excl;
			cbase[2 * lane + 1] = excl + cnt[0];
			if(lane == 63) {
				cbase[128] = incl;
				nkept = 128;
			}
		}
		__syncthreads();
		const int nk = nkept, M = cbase[nk];
		if(M > CELL_CMAX) {
			for(int h = tid / BUILD_LANES; h < nh; h += blockDim.x / BUILD_LANES)
				build_row<BUILD_LANES>(items[(size_t) c * cap + h], tid % BUILD_LANES, s, pos, bonds, systems, st, count, items, cap, nbr,
						nnbr, max_neigh, rlist, pos32, blist);
			__syncthreads();
			continue;
		}
		for(int m = tid; m < M; m += blockDim.x) {
			int lo = 0, hi = nk;   // the stencil cell holding candidate m
			while(hi - lo > 1) {
				const int mid = (lo + hi) >> 1;
				if(cbase[mid] <= m) lo = mid;
				else hi = mid;
			}
			const int j = items[(size_t) ccell[lo] * cap + (m - cbase[lo])];
			float4 q = (BACK && S.far_back) ? blist[j] : pos32[j];
			q.w = __int_as_float(j);
			cpos[m] = q;
		}
		__syncthreads();
		for(int h = w; h < nh; h += blockDim.x / 64) {
			const int i = items[(size_t) c * cap + h];
			const double4 p = pos[i];
			const float4 p32 = pos32[i];
			const int2 bd = bonds[i];
			const bool back = BACK && S.far_back;
			const float4 bi = back ? blist[i] : make_float4(0.f, 0.f, 0.f, 0.f);
			int *row = nbr + (size_t) i * max_neigh;
			int ncl = 0, nfa = 0;
			for(int m0 = 0; m0 < M; m0 += 64) {
				const int m = m0 + lane;
				bool close = false, far = false;
				int j = -1;
				if(m < M) {
					const float4 q = cpos[m];
					j = __float_as_int(q.w);
					if(j != i && j != bd.x && j != bd.y) {
						if(!back) classify_pair(S, p, p32, q, pos, j, true, close, far);
						else {
							const float db2 = back_dist2(bi, q, S);
							if(db2 < S.rgate2f) {
								classify_pair(S, p, p32, pos32[j], pos, j, true, close, far);
								if(far && !(db2 < S.rbf2)) far = false;
							}
						}
					}
				}
				const uint64_t bc = __ballot(close), bf = __ballot(far);
				if(close) {
					const int k = ncl + __popcll(bc & below);
					if(k + nfa < max_neigh) row[k] = j;
				}
				if(far) {
					const int k = nfa + __popcll(bf & below);
					if(ncl + k < max_neigh) row[max_neigh - 1 - k] = j;
				}
				ncl += __popcll(bc);
				nfa += __popcll(bf);
			}
			if(lane == 0) {
				if(ncl + nfa > max_neigh) {
					atomicOr(&st[s].error, 2);
					ncl = nfa = 0;
				}
				nnbr[i] = make_int2(ncl, nfa);
				if(rlist) rlist[i] = p;
			}
		}
		__syncthreads();
	}
}

// src = b.pos (the usual build) or b.rlist (the lists of the last build,
// rebuilt exactly, e.g. after a capacity change); rlist is then left as is
void launch_build_list(hipStream_t st, const Batch &b, int force_all, const double4 *src) {
	int bs = 256;
	size_t threads = (size_t) b.N * BUILD_LANES;
	const unsigned grid = (unsigned) std::min<size_t>((threads + bs - 1) / bs, 2048);
	if(b.cell_build && b.blist)
		k_build_cells<true><<<(unsigned) std::min(b.ncell_total, 4096), 256, 0, st>>>(b.ncell_total, src, b.bonds, b.d_sys, b.d_state,
				b.cell_owner, b.cell_count, b.cell_items, b.cell_cap, b.nbr, b.nnbr, b.max_neigh, src == b.rlist ? nullptr : b.rlist,
				force_all, b.pos32, b.any_rebuild, b.blist);
	else if(b.cell_build)
		k_build_cells<false><<<(unsigned) std::min(b.ncell_total, 4096), 256, 0, st>>>(b.ncell_total, src, b.bonds, b.d_sys, b.d_state,
				b.cell_owner, b.cell_count, b.cell_items, b.cell_cap, b.nbr, b.nnbr, b.max_neigh, src == b.rlist ? nullptr : b.rlist,
				force_all, b.pos32, b.any_rebuild, nullptr);
	else
		k_build_list<<<grid, bs, 0, st>>>(b.N, src, b.bonds, b.pinfo, b.d_sys, b.d_state, b.cell_count,
				b.cell_items, b.cell_cap, b.nbr, b.nnbr, b.max_neigh, src == b.rlist ? nullptr : b.rlist, force_all, b.pos32, b.any_rebuild,
				b.blist);
	if(b.ecap > 0)
		k_edge_rows<<<grid, bs, 0, st>>>(b.N, b.pinfo, b.d_sys, b.d_state, b.nbr, b.nnbr, b.max_neigh, b.ecap, b.eown, b.ein, b.ecount,
				force_all, b.any_rebuild);
}

// ------------------------------------------------------------ Hilbert sort

__device__ __forceinline__ unsigned int hilbert_key(unsigned int x, unsigned int y, unsigned int z, int bits) {
	unsigned int X[3] = {x, y, z};
	unsigned int M = 1u << (bits - 1);
	for(unsigned int Q = M; Q > 1; Q >>= 1) {
		unsigned int P = Q - 1;
#pragma unroll
		for(int i = 0; i < 3; i++) {
			if(X[i] & Q) X[0] ^= P;
			else {
				unsigned int t = (X[0] ^ X[i]) & P;
				X[0] ^= t;
				X[i] ^= t;
			}
		}
	}
	X[1] ^= X[0];
	X[2] ^= X[1];
	unsigned int t = 0;
	for(unsigned int Q = M; Q > 1; Q >>= 1)
		if(X[2] & Q) t ^= Q - 1;
	X[0] ^= t;
	X[1] ^= t;
	X[2] ^= t;
	unsigned int key = 0;
	for(int bb = bits - 1; bb >= 0; bb--) {
		key = (key << 3) | (((X[0] >> bb) & 1u) << 2) | (((X[1] >> bb) & 1u) << 1) | ((X[2] >> bb) & 1u);
	}
	return key;
}

__global__ void k_hilbert_keys(int N, const double4 *__restrict__ pos, const int2 *__restrict__ pinfo,
		const oxh_sys *__restrict__ systems, int bits, unsigned int *__restrict__ keys, int *__restrict__ vals) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= N) return;
	int s = pinfo[i].y;
	const oxh_sys &S = systems[s];
	double4 p = pos[i];
	int n = 1 << bits;
	unsigned int cx = cell_coord(p.x, S.inv_box[0], n);
	unsigned int cy = cell_coord(p.y, S.inv_box[1], n);
	unsigned int cz = cell_coord(p.z, S.inv_box[2], n);
	keys[i] = ((unsigned int) s << (3 * bits)) | hilbert_key(cx, cy, cz, bits);
	vals[i] = i;
}

// exported for the C ABI's key check (same function as the sort uses)
__global__ void k_hilbert_probe(int n, const uint4 *__restrict__ xyzb, unsigned int *__restrict__ out) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n) return;
	uint4 v = xyzb[i];
	out[i] = hilbert_key(v.x, v.y, v.z, (int) v.w);
}

constexpr int RADIX_TILE = 4096;   // keys per block and pass
constexpr int RADIX_BS = 256;

__global__ void __launch_bounds__(RADIX_BS) k_radix_hist(int N, const unsigned int *__restrict__ keys, int shift,
		unsigned int *__restrict__ hist, int nblocks) {
	__shared__ unsigned int h[256];
	h[threadIdx.x] = 0;
	__syncthreads();
	int base = blockIdx.x * RADIX_TILE;
	for(int k = threadIdx.x; k < RADIX_TILE; k += RADIX_BS) {
		int i = base + k;
		if(i < N) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
	}
	__syncthreads();
	hist[(size_t) threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of n entries in one workgroup (n is small: 256 * blocks)
__global__ void __launch_bounds__(1024) k_scan_single(unsigned int *__restrict__ data, int n) {
	__shared__ unsigned int part[1024];
	int t = threadIdx.x;
	int per = (n + 1023) / 1024;
	int b = t * per, e = min(n, b + per);
	unsigned int s = 0;
	for(int k = b; k < e; k++) s += data[k];
	part[t] = s;
	__syncthreads();
	for(int off = 1; off < 1024; off <<= 1) {
		unsigned int v = (t >= off) ? part[t - off] : 0u;
		__syncthreads();
		part[t] += v;
		__syncthreads();
	}
	unsigned int run = (t > 0) ? part[t - 1] : 0u;
	for(int k = b; k < e; k++) {
		unsigned int v = data[k];
		data[k] = run;
		run += v;
	}
}

__global__ void __launch_bounds__(RADIX_BS) k_radix_scatter(int N, const unsigned int *__restrict__ keys, const int *__restrict__ vals,
		int shift, const unsigned int *__restrict__ offs, int nblocks, unsigned int *__restrict__ keys_out, int *__restrict__ vals_out) {
	__shared__ unsigned int run[256];
	__shared__ unsigned int wcount[RADIX_BS / 64][256];
	const int t = threadIdx.x, w = t >> 6, lane = t & 63;
	run[t] = offs[(size_t) t * nblocks + blockIdx.x];
	for(int k = 0; k < RADIX_BS / 64; k++) wcount[k][t] = 0;
	__syncthreads();
	const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
	int base = blockIdx.x * RADIX_TILE;
	for(int c = 0; c < RADIX_TILE; c += RADIX_BS) {
		int i = base + c + t;
		bool valid = i < N;
		unsigned int key = valid ? keys[i] : 0u;
		unsigned int d = (key >> shift) & 255u;
		uint64_t m = __ballot(valid);
#pragma unroll
		for(int bit = 0; bit < 8; bit++) {
			uint64_t bal = __ballot(((d >> bit) & 1u) != 0u);
			m &= ((d >> bit) & 1u) ? bal : ~bal;
		}
		unsigned int rank = __popcll(m & lt);
		if(valid && rank == 0) wcount[w][d] = __popcll(m);
		__syncthreads();
		if(valid) {
			unsigned int off = run[d] + rank;
			for(int k = 0; k < w; k++) off += wcount[k][d];
			keys_out[off] = key;
			vals_out[off] = vals[i];
		}
		__syncthreads();
		unsigned int tot = 0;
		for(int k = 0; k < RADIX_BS / 64; k++) {
			tot += wcount[k][t];
			wcount[k][t] = 0;
		}
		run[t] += tot;
		__syncthreads();
	}
}

template<typename T>
__global__ void k_gather(int N, const int *__restrict__ perm, const T *__restrict__ src, T *__restrict__ dst) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i < N) dst[i] = src[perm[i]];
}

__global__ void k_invert(int N, const int *__restrict__ perm, int *__restrict__ inv) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i < N) inv[perm[i]] = i;
}

__global__ void k_gather_bonds(int N, const int *__restrict__ perm, const int *__restrict__ inv, const int2 *__restrict__ src,
		int2 *__restrict__ dst) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= N) return;
	int2 b = src[perm[i]];
	dst[i] = make_int2(b.x >= 0 ? inv[b.x] : -1, b.y >= 0 ? inv[b.y] : -1);
}

__global__ void k_mark_rebuild(int nsys, oxh_sys_state *st) {
	int s = blockIdx.x * blockDim.x + threadIdx.x;
	if(s < nsys) st[s].need_rebuild = 1;
}

// Reorders every system's particles along a Hilbert curve of its box.
// Returns 0 on success, -1 if the batch has too many systems to encode.
int launch_hilbert_sort(hipStream_t st, Batch &b) {
	int N = b.N;
	if(N < 2) return 0;
	int sys_bits = 0;
	while((1 << sys_bits) < b.nsys) sys_bits++;
	int bits = (32 - sys_bits) / 3;
	if(bits > 10) bits = 10;
	if(bits < 2) return -1;
	int total_bits = 3 * bits + sys_bits;
	int bs = 256, g = (N + bs - 1) / bs;
	k_hilbert_keys<<<g, bs, 0, st>>>(N, b.pos, b.pinfo, b.d_sys, bits, b.keys, b.perm);
	int nblocks = (N + RADIX_TILE - 1) / RADIX_TILE;
	unsigned int *kin = b.keys, *kout = b.keys_alt;
	int *vin = b.perm, *vout = b.perm_alt;
	for(int shift = 0; shift < total_bits; shift += 8) {
		k_radix_hist<<<nblocks, RADIX_BS, 0, st>>>(N, kin, shift, b.radix_hist, nblocks);
		k_scan_single<<<1, 1024, 0, st>>>(b.radix_hist, 256 * nblocks);
		k_radix_scatter<<<nblocks, RADIX_BS, 0, st>>>(N, kin, vin, shift, b.radix_hist, nblocks, kout, vout);
		std::swap(kin, kout);
		std::swap(vin, vout);
	}
	const int *perm = vin;
	k_invert<<<g, bs, 0, st>>>(N, perm, b.inv_perm);
	// gather every per-particle array through the scratch buffer
	auto permute = [&](auto *arr) {
		using T = std::remove_pointer_t<decltype(arr)>;
		T *tmp = (T *) b.scratch;
		k_gather<T><<<g, bs, 0, st>>>(N, perm, arr, tmp);
		(void) hipMemcpyAsync(arr, tmp, sizeof(T) * N, hipMemcpyDeviceToDevice, st);
	};
	permute(b.pos);
	permute(b.vel);
	permute(b.Lm);
	permute(b.q64);
	permute(b.F);
	permute(b.T);
	if(b.fp64) {
		permute(b.F64);
		permute(b.T64);
	}
	permute(b.pinfo);
	permute(b.orig);
	permute(b.pstep);
	permute(b.img);
	{
		int2 *tmp = (int2 *) b.scratch;
		k_gather_bonds<<<g, bs, 0, st>>>(N, perm, b.inv_perm, b.bonds, tmp);
		(void) hipMemcpyAsync(b.bonds, tmp, sizeof(int2) * N, hipMemcpyDeviceToDevice, st);
	}
	launch_prepare_mirror(st, b);
	launch_slot_of(st, b);
	k_mark_rebuild<<<(b.nsys + 63) / 64, 64, 0, st>>>(b.nsys, b.d_state);
	return 0;
}

}  // namespace oxh

// key check entry used by the C ABI (no allocation inside: caller buffers)
extern "C" int oxh_internal_hilbert_probe(int n, const void *d_xyzb, void *d_out, hipStream_t st) {
	if(n <= 0) return 0;
	oxh::k_hilbert_probe<<<(n + 255) / 256, 256, 0, st>>>(n, (const uint4 *) d_xyzb, (unsigned int *) d_out);
	return (int) hipGetLastError();
}
