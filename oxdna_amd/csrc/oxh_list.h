// Verlet-list construction of one particle on gfx950, shared by the
// one-kernel-per-phase path (oxh_lists.hip) and the workgroup-resident
// engine (oxh_persist.hip).
#pragma once

#include <type_traits>

#include "oxh_device.h"

namespace oxh {

__device__ __forceinline__ int cell_coord(double x, double inv_box, int nc) {
	double s = x * inv_box;
	double f = s - floor(s);
	int c = (int) (f * nc);
	return c < nc ? c : nc - 1;
}

// FP64 squared nearest-image distance, fixed operation sequence
__device__ __forceinline__ double sqr_dist(double4 a, double4 b, const oxh_sys &S) {
	double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
	dx = fma(-S.box[0], rint(dx * S.inv_box[0]), dx);
	dy = fma(-S.box[1], rint(dy * S.inv_box[1]), dy);
	dz = fma(-S.box[2], rint(dz * S.inv_box[2]), dz);
	return fma(dz, dz, fma(dy, dy, dx * dx));
}

constexpr int BUILD_LANES = 32;

// Far pairs kept by their backbone sites (S.far_back): a far pair enters
// the list only while the backbone sites recorded at the build (blist:
// wrapped FP32 sites) are within sqrt(rbf2).  Far pairs interact through
// Debye-Hueckel alone, between backbone sites and within its cut-off; rbf2
// adds twice the largest site move allowed before a rebuild, so a pair
// left out stays beyond the cut-off, with a zero term, until the next
// build.  The sites' distance is taken first: beyond sqrt(rgate2f) (which
// no close pair's sites reach) a candidate is no pair, and only the rest
// have their centres compared.
__device__ __forceinline__ float back_dist2(float4 a, float4 b, const oxh_sys &S) {
	float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
	dx -= S.boxf[0] * rintf(dx * S.inv_boxf[0]);
	dy -= S.boxf[1] * rintf(dy * S.inv_boxf[1]);
	dz -= S.boxf[2] * rintf(dz * S.inv_boxf[2]);
	return dx * dx + dy * dy + dz * dz;
}

// The pair predicate of the lists: candidate j of particle i (FP64 centres
// p, pj; FP32 wrapped centres p32, q for the prefilter when from_cells) is a
// close pair (closer than sqrt(rclose2)) or a far pair (closer than r_verlet).
// FP32 decides unless within its error bound of a threshold, FP64 then.
__device__ __forceinline__ void classify_pair(const oxh_sys &S, double4 p, float4 p32, float4 q, const double4 *__restrict__ pos,
		int j, bool from_cells, bool &close, bool &far) {
	close = far = false;
	if(from_cells) {
		float dx = q.x - p32.x, dy = q.y - p32.y, dz = q.z - p32.z;
		dx -= S.boxf[0] * rintf(dx * S.inv_boxf[0]);
		dy -= S.boxf[1] * rintf(dy * S.inv_boxf[1]);
		dz -= S.boxf[2] * rintf(dz * S.inv_boxf[2]);
		const float r2f = dx * dx + dy * dy + dz * dz;
		if(fabsf(r2f - S.rv2f) < S.eps2 || fabsf(r2f - S.rclose2f) < S.eps2) {
			double r2 = sqr_dist(p, pos[j], S);
			if(r2 < S.rv2) {
				close = r2 < S.rclose2;
				far = !close;
			}
		}
		else if(r2f < S.rv2f) {
			close = r2f < S.rclose2f;
			far = !close;
		}
	}
	else {
		double r2 = sqr_dist(p, pos[j], S);
		if(r2 < S.rv2) {
			close = r2 < S.rclose2;
			far = !close;
		}
	}
}

// BUILD_LANES threads per particle share the candidate scan: in every chunk
// each lane tests one candidate, a wave ballot gives every passing lane its
// slot, so the row is written in candidate order whatever the timing.
// Pairs closer than sqrt(rclose2) go to the front of the row, the rest to
// the back.  A pair cannot come closer than R_CLOSE before the next rebuild
// unless it is within R_CLOSE + 2 skin now, so far pairs only ever need
// Debye-Hueckel.
// With cells, the candidates are the particles of the cells within
// +-stencil of the particle's cell (27 or 125 cells; k = (oz + R) W^2 +
// (oy + R) W + ox + R), minus the cells whose nearest point is out of
// range.  With K = 32 lanes, lane l takes cells l, l + 32, ...; lane prefix
// sums of the cells' occupancies number the candidates (cells in k order,
// each cell's items sorted by particle index), and the group scans them in
// chunks of K, each lane finding its candidate's cell by a binary search
// over the prefix sums: the cells' loads overlap instead of following each
// other.
// Every lane of the group must call it; lane 0 writes the row's counts and
// the list position.
template<int K>
__device__ __forceinline__ void build_row(int i, int l, int s, const double4 *__restrict__ pos, const int2 *__restrict__ bonds,
		const oxh_sys *__restrict__ systems, oxh_sys_state *__restrict__ st, const int *__restrict__ count,
		const int *__restrict__ items, int cap, int *__restrict__ nbr, int2 *__restrict__ nnbr, int max_neigh,
		double4 *__restrict__ rlist, const float4 *__restrict__ pos32, const float4 *__restrict__ blist) {
	const oxh_sys &S = systems[s];
	const double4 p = pos[i];
	const int2 bd = bonds[i];
	const bool back = S.far_back && blist;
	int *row = nbr + (size_t) i * max_neigh;
	const float4 p32 = S.use_cells ? pos32[i] : make_float4(0.f, 0.f, 0.f, 0.f);
	const float4 bi = back ? blist[i] : make_float4(0.f, 0.f, 0.f, 0.f);
	const int g0 = (threadIdx.x & 63) & ~(K - 1);          // first lane of this group in the wave
	const uint64_t gmask = (K == 64) ? ~0ull : (((1ull << K) - 1ull) << g0);
	const uint64_t below = ((1ull << (threadIdx.x & 63)) - 1ull) & gmask;
	int nc = 0, nf = 0;
	// one chunk: lane tests candidate j (-1: none); from_cells selects the
	// FP32 prefilter on wrapped centres
	auto test = [&](int j, bool from_cells) {
		bool close = false, far = false;
		if(j >= 0 && j != i && j != bd.x && j != bd.y) {
			const float db2 = back ? back_dist2(bi, blist[j], S) : 0.f;
			if(!back || db2 < S.rgate2f) {
				classify_pair(S, p, p32, from_cells ? pos32[j] : make_float4(0.f, 0.f, 0.f, 0.f), pos, j, from_cells, close, far);
				if(far && back && !(db2 < S.rbf2)) far = false;
			}
		}
		const uint64_t bc = __ballot(close) & gmask, bf = __ballot(far) & gmask;
		if(close) {
			int k = nc + __popcll(bc & below);
			if(k + nf < max_neigh) row[k] = j;
		}
		if(far) {
			int k = nf + __popcll(bf & below);
			if(nc + k < max_neigh) row[max_neigh - 1 - k] = j;
		}
		nc += __popcll(bc);
		nf += __popcll(bf);
	};
	if(S.use_cells) {
		const int cx = cell_coord(p.x, S.inv_box[0], S.nc[0]);
		const int cy = cell_coord(p.y, S.inv_box[1], S.nc[1]);
		const int cz = cell_coord(p.z, S.inv_box[2], S.nc[2]);
		const int R = S.stencil, W = 2 * R + 1, NS = W * W * W;
		// position inside its cell, for the distance to other cells
		const float ux = p32.x - cx * S.cell_w[0], uy = p32.y - cy * S.cell_w[1], uz = p32.z - cz * S.cell_w[2];
		// a cell is culled when its nearest point is farther than r_verlet
		// by a margin far above the FP32 error of these distances
		const float cull = (sqrtf(S.rv2f) + 1e-3f) * (sqrtf(S.rv2f) + 1e-3f);
		auto gap = [](int o, float u, float w) { return o > 0 ? o * w - u : (o < 0 ? u + (-o - 1) * w : 0.f); };
		auto cell_of = [&](int k, bool &in_range) {
			const int oz = k / (W * W) - R, oy = (k / W) % W - R, ox = k % W - R;
			const float gx = gap(ox, ux, S.cell_w[0]), gy = gap(oy, uy, S.cell_w[1]), gz = gap(oz, uz, S.cell_w[2]);
			in_range = gx * gx + gy * gy + gz * gz <= cull;
			int z = cz + oz, y = cy + oy, x = cx + ox;
			z += (z < 0) ? S.nc[2] : (z >= S.nc[2] ? -S.nc[2] : 0);
			y += (y < 0) ? S.nc[1] : (y >= S.nc[1] ? -S.nc[1] : 0);
			x += (x < 0) ? S.nc[0] : (x >= S.nc[0] ? -S.nc[0] : 0);
			return S.cell_first + (z * S.nc[1] + y) * S.nc[0] + x;
		};
		if constexpr(K == 32) {
			// lane l takes cells l, l + 32, ... (1 round for 27 cells, 4 for
			// 125); the stencil is a compile-time constant of each branch
			auto stream = [&](auto rounds_tag) {
				constexpr int ROUNDS = decltype(rounds_tag)::value;
				int cell[ROUNDS], excl[ROUNDS], base[ROUNDS + 1];
				base[0] = 0;
#pragma unroll
				for(int r = 0; r < ROUNDS; r++) {
					const int k = l + K * r;
					int cnt = 0;
					cell[r] = 0;
					if(k < NS) {
						bool in_range;
						cell[r] = cell_of(k, in_range);
						if(in_range) cnt = min(count[cell[r]], cap);
					}
					int incl = cnt;
					for(int off = 1; off < K; off <<= 1) {
						const int v = __shfl_up(incl, off, K);
						if(l >= off) incl += v;
					}
					excl[r] = incl - cnt;
					base[r + 1] = base[r] + __shfl(incl, K - 1, K);
				}
				const int total = base[ROUNDS];
				for(int m0 = 0; m0 < total; m0 += K) {
					const int g = m0 + l;
					int c = 0, slot = 0;
#pragma unroll
					for(int r = 0; r < ROUNDS; r++) {
						// candidates base[r] .. base[r + 1] - 1 come from round r
						const int gr = g - base[r];
						int lo = 0;
						for(int step = K >> 1; step > 0; step >>= 1)
							if(__shfl(excl[r], lo + step, K) <= gr) lo += step;
						const int cr = __shfl(cell[r], lo, K), sr = gr - __shfl(excl[r], lo, K);
						if(g >= base[r] && g < base[r + 1]) {
							c = cr;
							slot = sr;
						}
					}
					test(g < total ? items[(size_t) c * cap + slot] : -1, true);
				}
			};
			if(R == 1) stream(std::integral_constant<int, 1>());
			else stream(std::integral_constant<int, 4>());
		}
		else {
			for(int k = 0; k < NS; k++) {
				bool in_range;
				const int c = cell_of(k, in_range);
				const int n = in_range ? min(count[c], cap) : 0;
				for(int m0 = 0; m0 < n; m0 += K) test(m0 + l < n ? items[(size_t) c * cap + m0 + l] : -1, true);
			}
		}
	}
	else {
		for(int m0 = 0; m0 < S.n; m0 += K) test(m0 + l < S.n ? S.first + m0 + l : -1, false);
	}
	if(l != 0) return;
	if(nc + nf > max_neigh) {
		atomicOr(&st[s].error, 2);
		nc = nf = 0;
	}
	nnbr[i] = make_int2(nc, nf);
	if(rlist) rlist[i] = p;
}

}  // namespace oxh
