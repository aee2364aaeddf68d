// Verlet-list construction of one particle on gfx950, shared by the
// one-kernel-per-phase path (oxh_lists.hip) and the workgroup-resident
// engine (oxh_persist.hip).
#pragma once

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

constexpr int BUILD_LANES = 16;

// BUILD_LANES threads per particle share the candidate scan: in every chunk
// each lane tests one candidate, a wave ballot gives every passing lane its
// slot, so the row is written in candidate order whatever the timing.
// Pairs closer than sqrt(rclose2) go to the front of the row, the rest to
// the back.  A pair cannot come closer than R_CLOSE before the next rebuild
// unless it is within R_CLOSE + 2 skin now, so far pairs only ever need
// Debye-Hueckel.
// Every lane of the group must call it; lane 0 writes the row's counts and
// the list position.
template<int K>
__device__ __forceinline__ void build_row(int i, int l, int s, const double4 *__restrict__ pos, const int2 *__restrict__ bonds,
		const oxh_sys *__restrict__ systems, oxh_sys_state *__restrict__ st, const int *__restrict__ count,
		const int *__restrict__ items, int cap, int *__restrict__ nbr, int2 *__restrict__ nnbr, int max_neigh,
		double4 *__restrict__ rlist, const float4 *__restrict__ pos32) {
	const oxh_sys &S = systems[s];
	const double4 p = pos[i];
	const int2 bd = bonds[i];
	int *row = nbr + (size_t) i * max_neigh;
	const float4 p32 = S.use_cells ? pos32[i] : make_float4(0.f, 0.f, 0.f, 0.f);
	const int g0 = (threadIdx.x & 63) & ~(K - 1);          // first lane of this group in the wave
	const uint64_t gmask = (K == 64) ? ~0ull : (((1ull << K) - 1ull) << g0);
	const uint64_t below = ((1ull << (threadIdx.x & 63)) - 1ull) & gmask;
	int nc = 0, nf = 0;
	// tests candidates base[0 .. n) in chunks of K
	auto scan = [&](const int *cand, int first, int n) {
		for(int m0 = 0; m0 < n; m0 += K) {
			const int m = m0 + l;
			bool close = false, far = false;
			int j = -1;
			if(m < n) {
				j = cand ? cand[m] : first + m;
				if(j != i && j != bd.x && j != bd.y) {
					if(cand) {
						// FP32 decides unless within its error bound of a threshold
						const float4 q = pos32[j];
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
		}
	};
	if(S.use_cells) {
		int cx = cell_coord(p.x, S.inv_box[0], S.nc[0]);
		int cy = cell_coord(p.y, S.inv_box[1], S.nc[1]);
		int cz = cell_coord(p.z, S.inv_box[2], S.nc[2]);
		for(int dz = -1; dz <= 1; dz++) {
			int z = cz + dz;
			z += (z < 0) ? S.nc[2] : (z >= S.nc[2] ? -S.nc[2] : 0);
			for(int dy = -1; dy <= 1; dy++) {
				int y = cy + dy;
				y += (y < 0) ? S.nc[1] : (y >= S.nc[1] ? -S.nc[1] : 0);
				for(int dx = -1; dx <= 1; dx++) {
					int x = cx + dx;
					x += (x < 0) ? S.nc[0] : (x >= S.nc[0] ? -S.nc[0] : 0);
					int c = S.cell_first + (z * S.nc[1] + y) * S.nc[0] + x;
					scan(items + (size_t) c * cap, 0, min(count[c], cap));
				}
			}
		}
	}
	else {
		scan(nullptr, S.first, S.n);
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
