// Types shared by the host side and the gfx950 kernels.
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Per-system parameters, resident in HBM (one entry per system of a batch).
// Everything the kernels need about a system: where its particles live in
// the batch arrays, its box, its cells, its T-dependent model constants and
// its integrator / thermostat settings.
typedef struct oxh_sys {
	// geometry, FP64 (minimum image and wrapping are done in FP64)
	double box[3];
	double inv_box[3];
	float boxf[3], inv_boxf[3];   // FP32 copies for the far-pair nearest image
	// particle range [first, first + n) in the batch arrays
	int first, n;
	// cells: nc[3] per side, [cell_first, cell_first + ncell) in the batch's
	// cell arrays; use_cells = 0 selects the all-pairs list build
	int nc[3];
	int cell_first, ncell, use_cells;
	// neighbour-cell stencil: cells within +-stencil of a particle's cell
	// (1: cell edges >= r_verlet, 27 cells; 2: edges >= r_verlet / 2, 125
	// cells); cell edge lengths (FP32), for culling cells out of range
	int stencil;
	float cell_w[3];
	// 1 when a box side is shorter than 2 (r_verlet + 2 backbone arms): the
	// nearest image of two backbone sites may then differ from the image the
	// centres select, so far pairs take the centres' image
	int small_box;
	// Verlet list: r_verlet^2, skin^2 and the build-time bound below which a
	// pair goes to the close sublist, (R_CLOSE + 2 skin)^2, FP64
	double rv2, skin2, rclose2;
	// FP32 squared distances within eps2 of rv2 or rclose2 are re-decided in
	// FP64 (bound on the FP32 rounding error of wrapped coordinates)
	float rv2f, rclose2f, eps2;
	// far pairs kept by their backbone sites (far_back = 1): a far pair
	// enters the list only while its backbone sites at the build are
	// closer than sqrt(rbf2), the Debye-Hueckel cut-off plus twice the
	// largest backbone-site move allowed before a rebuild plus a margin
	// above the FP32 error; a site that moves more than sqrt(sback2f) from
	// its place at the build triggers a rebuild.  Candidates whose sites
	// are farther apart than sqrt(rgate2f) (at least sqrt(rbf2), and beyond
	// the sites' reach of any close pair) are no pair at all.
	int far_back;
	float rbf2, sback2f, rgate2f;
	// model constants (FP32) that depend on T / salt / options
	float f1_eps[2][5][5];
	float f1_shift[2][5][5];
	float dh_prefactor, dh_minus_kappa, dh_rhigh, dh_rc, dh_b;
	int dh_half;
	float hb_multi;
	int use_mbf;
	float mbf_xmax, mbf_fmax, mbf_finf;
	float fene_r0;
	// the same constants in FP64, for the all-FP64 force mode
	double f1_eps_d[2][5][5];
	double f1_shift_d[2][5][5];
	double dh_prefactor_d, dh_minus_kappa_d, dh_rhigh_d, dh_rc_d, dh_b_d;
	double hb_multi_d, mbf_xmax_d, mbf_fmax_d, mbf_finf_d, fene_r0_d;
	// integration, FP64
	double dt, T;
	int thermostat;       // 0 none, 1 Langevin, 2 Bussi, 3 Brownian
	int newtonian_steps;
	double gamma_t, resc_t, gamma_r, resc_r;   // Langevin
	double pt, pr, resc_brown;                  // Brownian
	double bussi_c;                             // exp(-newtonian_steps / tau)
	uint32_t seed_lo, seed_hi;
	int pad_;
	// external forces: the reference's step counter at this system's first
	// step (forces of step k use ext_step0 + k), and 1 when any record acts
	int has_ext;
	long long ext_step0;
} oxh_sys;

// One external force acting on one particle (the reference's
// external_forces_file entries, src/Forces/*.cpp, expanded to one record
// per particle and stored sorted by particle, in the file's order).
// Parameters by type (OXH_EXT_* in include/oxh.h; unit = normalised):
//   string:          a = F0, b = rate, v1 = dir (unit, or the centre when
//                    flags & 2: dir_as_centre)
//   trap:            a = stiff, b = rate, v1 = dir (unit), v2 = pos0
//   lowdim_trap:     as trap; flags bits 0-2 = visibility of x, y, z
//   mutual_trap:     a = stiff, b = stiff_rate, c = r0, d = rate, ref
//   constant_trap:   a = stiff, c = r0, ref
//   repulsion_plane: a = stiff, b = position, c = end_position, d = v,
//                    v1 = dir (unit)
//   attraction_plane: a = stiff, b = position, v1 = dir (unit)
//   repulsion_plane_moving: a = stiff, v1 = dir (unit), reference
//                    particles ref..ref2
//   sphere:          a = stiff, b = rate, c = r0, d = r_ext, v1 = center
//   sphere_smooth:   a = stiff, c = r0, d = r_ext, e = smooth, f = alpha,
//                    v1 = center
//   repulsive_sphere_moving: a = stiff, b = rate, c = r0, d = r_ext,
//                    v1 = origin, v2 = target, steps
//   repulsive_kepler_poinsot: a = stiff, b = rate, c = apex, d = base,
//                    e = base_radius, f = kappa, v1 = center
//   LJ_wall:         a = stiff, b = position, c = sigma, d = cutoff, n,
//                    v1 = dir (unit)
//   LJ_cone:         a = stiff, c = sigma, d = cutoff, e/f/g = sin/cos/tan
//                    of alpha, n, v1 = dir (unit), v2 = pos0
//   twist:           a = stiff, b = rate, c = base, v1 = axis (unit),
//                    v2 = pos0, v3 = center, v4 = mask
//   generic_central_force: a = F0, c = inner cut-off, d = outer cut-off,
//                    v1 = center
//   ellipsoid:       a = stiff, v1 = center, v2 = r_1, v3 = r_2
//   com:             a = stiff, b = rate, c = r0, ref = centre group,
//                    n = particles in the com list
//   yukawa_sphere:   a = radius, b = WCA epsilon, c = WCA sigma, d = Debye
//                    length, e = Debye amplitude, f = cutoff, g = WCA
//                    cutoff, n = WCA n, v1 = center
// ref is the partner's index inside the system (upload order); flags & 1:
// nearest image between the two particles (PBC).
typedef struct oxh_ext_rec {
	int type, flags, ref, ref2;   // ref2: the last of a range of reference particles
	int n, pad_;
	long long steps;
	double a, b, c, d, e, f, g;
	double v1[3], v2[3], v3[3], v4[3];
} oxh_ext_rec;

// per-system device-side state written by the kernels
typedef struct oxh_sys_state {
	double K_t, K_r;          // Bussi thermostat internal kinetic energies
	int need_rebuild;         // skin trigger
	int error;                // bit 0: FENE bond broken, bit 1: list overflow, bit 2: cell overflow, bit 3: edge-row overflow
	long long n_rebuilds;
} oxh_sys_state;

#ifdef __cplusplus
}
#endif
