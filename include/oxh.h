/*
 * oxh: MI355X-native oxDNA2 molecular-dynamics hot path, C ABI.
 *
 * Plain C types only (pointers, sizes, scalars): no torch, no HIP types.
 * Every call returns an oxh_status; 0 is success, and oxh_last_error()
 * describes the last failure of the calling thread.  Nothing in the
 * library calls exit(): a bad system is reported, not fatal (the
 * reference exits on any CUDA error, src/CUDA/cuda_defs.h:12-22).
 *
 * A batch holds one or more independent systems resident in HBM.  The
 * reference runs one system per process (src/Utilities/ConfigInfo.h:47,
 * 94-98); the plain path here is a batch of one system, the batched and
 * served forms are batches of many.
 *
 * Threading: thread-compatible, not thread-safe per batch handle.
 */
#ifndef OXH_H
#define OXH_H

#ifdef __cplusplus
extern "C" {
#endif

typedef int oxh_status;

#define OXH_OK 0
#define OXH_ERR_ARG 1        /* invalid argument */
#define OXH_ERR_HIP 2        /* HIP runtime failure */
#define OXH_ERR_ALLOC 3      /* device allocation failed */
#define OXH_ERR_SYSTEM 4     /* a system failed (broken bond, overflow); see oxh_batch_system_error */

typedef struct oxh_batch oxh_batch;

/* Thermostat codes, the reference's `thermostat` input key
 * (src/Backends/Thermostats/ThermostatFactory.cpp): */
#define OXH_THERMOSTAT_NO 0
#define OXH_THERMOSTAT_LANGEVIN 1
#define OXH_THERMOSTAT_BUSSI 2
#define OXH_THERMOSTAT_BROWNIAN 3   /* "brownian" / "john" */

/* interaction_type (src/Interactions/InteractionFactory.cpp:51-89) */
#define OXH_INTERACTION_DNA2 0      /* "DNA2" / "DNA2_nomesh": oxDNA2 (DNA2Interaction) */
#define OXH_INTERACTION_RNA2 1      /* "RNA2": oxRNA2, oxRNA with Debye-Hueckel (RNA2Interaction) */
#define OXH_INTERACTION_RNA 2       /* "RNA": oxRNA (RNAInteraction) */

/* One system: state, topology and the input-file keys that set up the
 * interaction, the list and the integrator.  Arrays are host arrays in the
 * system's own particle order; they are copied at upload.
 * Keys and defaults follow the reference (docs/source/input.md:29-113):
 *   T (simulation units; see src/Utilities/Utils.cpp:316-352 for "20C"),
 *   salt_concentration, dh_lambda (0.3616455), dh_strength (0.0543),
 *   debye_huckel_rhigh (3 lambda), dh_half_charged_ends (1),
 *   hb_multiplier (1), max_backbone_force[_far] (off / 0.04), verlet_skin,
 *   dt, thermostat, diff_coeff or gamma_trans, newtonian_steps, pt,
 *   bussi_tau, seed, refresh_vel. */
typedef struct oxh_system_desc {
	int N;
	const double *pos;       /* 3N */
	const double *a1;        /* 3N, principal axis */
	const double *a3;        /* 3N, stacking axis (a2 = a3 x a1) */
	const double *vel;       /* 3N or NULL (zero, or refreshed) */
	const double *L;         /* 3N or NULL, body frame angular momentum */
	const int *btype;        /* N: A=0 G=1 C=2 T=3, or custom base types */
	const int *n3;           /* N: 3' neighbour, -1 = none */
	const int *n5;           /* N: 5' neighbour, -1 = none */
	double box[3];

	double T;
	double salt;
	int dh_half_charged_ends;
	double dh_lambda;        /* < 0: default */
	double dh_strength;      /* < 0: default */
	double dh_rhigh;         /* <= 0: default */
	double hb_multiplier;
	double max_backbone_force;      /* < 0: off */
	double max_backbone_force_far;  /* < 0: default 0.04 */
	/* sequence-dependent well depths (use_average_seq = false): NULL keeps
	 * the average-sequence oxDNA2 values.  stck[4][4] for A,G,C,T. */
	const double *seq_stck;
	double seq_stck_fact_eps;
	double seq_hydr_AT, seq_hydr_GC;

	double verlet_skin;
	double dt;
	int thermostat;
	double diff_coeff;       /* Langevin / Brownian; <= 0: use gamma_trans / pt */
	double gamma_trans;
	double pt;               /* Brownian; <= 0: from diff_coeff */
	int newtonian_steps;
	int bussi_tau;
	unsigned long long seed;
	int refresh_vel;         /* 1: draw vel and L from N(0, T) at upload */
	/* the reference's backend_precision key (src/Backends/BackendFactory.cpp:
	 * 48-80): 0 "mixed" (FP32 pair terms from FP64 site geometry, FP64
	 * state; default), 1 "double" (every term in FP64).  All systems of a
	 * batch use the same; slot pools take mixed only. */
	int backend_precision;
	/* OXH_INTERACTION_*: the model of the pair terms.  All systems of a
	 * batch use the same.  The oxRNA models take the average-sequence
	 * parameters (no seq_stck), salt / dh_* / dh_half_charged_ends /
	 * max_backbone_force as RNA2Interaction::get_settings reads them
	 * (src/Interactions/RNAInteraction2.cpp:33-72; dh_lambda default
	 * 0.3667258, dh_strength 0.0858), and run on the grid engine
	 * (not in slot pools or the workgroup engine). */
	int interaction;
} oxh_system_desc;

/* Fills a descriptor with the reference's defaults (no arrays set). */
void oxh_system_desc_defaults(oxh_system_desc *d);

const char *oxh_last_error(void);
int oxh_version(void);
oxh_status oxh_device_count(int *n);

/* Uploads n_sys systems to device `device` and builds their lists and
 * forces.  Replaces the reference's backend set-up sequence
 * MD_CUDABackend::init (src/CUDA/Backends/MD_CUDABackend.cu:674-755):
 * CUDABaseBackend::init_cuda, apply_changes_to_simulation_data (:231-313),
 * CUDASimpleVerletList::init (src/CUDA/Lists/CUDASimpleVerletList.cu:
 * 165-202), thermostat init, first list build + compute_forces. */
oxh_status oxh_batch_create(int device, int n_sys, const oxh_system_desc *systems, oxh_batch **out);
oxh_status oxh_batch_destroy(oxh_batch *b);

/* Rebuilds the Verlet lists: of every system when force != 0, else only of
 * systems whose particles moved more than the skin.  Replaces
 * CUDABaseList::update / CUDASimpleVerletList::update
 * (src/CUDA/Lists/CUDABaseList.h:55, CUDASimpleVerletList.cu:231-289). */
oxh_status oxh_batch_update_lists(oxh_batch *b, int force);

/* Forces and torques of the batch's current state into the batch's own
 * buffers (F lab frame, T body frame; written, not accumulated, with the
 * batch's external forces added; read them with oxh_batch_get_forces).
 * The kernels are CUDADNAInteraction's (src/CUDA/Interactions/
 * CUDA_DNA.cuh:832-904); oxh_interaction_compute_forces below is the
 * accumulating surface of CUDABaseInteraction::compute_forces. */
oxh_status oxh_batch_compute_forces(oxh_batch *b);

/* The GPU interaction surface with caller-owned device buffers, as
 * CUDABaseInteraction::compute_forces(lists, d_poss, d_orientations,
 * d_forces, d_torques, d_bonds, d_box) (src/CUDA/Interactions/
 * CUDABaseInteraction.h:60, called by MD_CUDABackend::_forces_second_step,
 * src/CUDA/Backends/MD_CUDABackend.cu:518-520, after _set_external_forces
 * (:528) has filled F and T): the pair forces of the state in d_pos / d_quat are
 * ADDED to what d_forces / d_torques already hold (external forces, or
 * zeros).  Batches only (not slot pools).
 *   d_pos     4 doubles per particle (x, y, z, unused), absolute positions;
 *   d_quat    4 doubles per particle, the unit quaternion (x, y, z, w) of
 *             the rotation whose matrix has the columns a1, a2 = a3 x a1,
 *             a3 (the reference's p->orientation);
 *   d_forces  4 numbers per particle: force (lab frame) and, in .w, the
 *             particle's share of the potential energy;
 *   d_torques 4 numbers per particle: torque (body frame, as the
 *             reference's p->torque) and, in .w, its share of the
 *             hydrogen-bonding energy.
 * The numbers are floats when the batch was created with
 * backend_precision = mixed and doubles with double (the reference's
 * c_number4).  Particle k of system s sits at index first_s + k, first_s =
 * N_0 + ... + N_{s-1}.  Topology, box and parameters are the batch's.  The
 * state becomes the batch's; the lists are rebuilt for the systems in which
 * a particle moved more than the Verlet skin since their last build (the
 * reference's CUDABaseList::update when _d_are_lists_old,
 * src/CUDA/Backends/MD_CUDABackend.cu:580-590).  stream: a hipStream_t, or
 * NULL for the batch's own; the additions are queued on it (the call
 * returns after the list check, before they finish). */
oxh_status oxh_interaction_compute_forces(oxh_batch *b, const double *d_pos, const double *d_quat, void *d_forces, void *d_torques,
		void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the same additions as oxh_interaction_compute_forces, for ONE
 * system described by desc (topology from its btype / n3 / n5 host arrays,
 * parameters and backend_precision from its keys; its pos, a1, a3, vel and
 * L are not read), with the state in d_pos / d_quat and the results added
 * onto d_forces / d_torques (layouts as above, particle k at index k).
 * Nothing is allocated: the Verlet list (built afresh on every call, all
 * pairs), the FP32 mirror, F, T and the parameters live in a fixed device
 * array of the library, OXH_NOALLOC_SCRATCH_BYTES on each GPU it runs on,
 * held until the process ends.  Limits: at most OXH_NOALLOC_MAX_N
 * nucleotides (larger N is rejected with OXH_ERR_ARG before any GPU work)
 * and OXH_NOALLOC_MAX_NEIGH Verlet neighbours per nucleotide (more returns
 * OXH_ERR_SYSTEM and adds nothing).  The call waits for the list and force
 * kernels, then queues the additions on `stream` (NULL: the null stream)
 * and returns.  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (other threads, other streams, a call issued before
 * the previous call's additions have run) race; serialise them. */
#define OXH_NOALLOC_MAX_N 4096
#define OXH_NOALLOC_MAX_NEIGH 256
#define OXH_NOALLOC_SCRATCH_BYTES 5425408
oxh_status oxh_interaction_compute_forces_noalloc(int device, const oxh_system_desc *desc, const double *d_pos, const double *d_quat,
		void *d_forces, void *d_torques, void *stream);

/* Advances every system by n_steps MD steps: first half kick + drift +
 * rotation, skin-triggered list rebuild, forces, second half kick,
 * thermostat.  Replaces the loop of MD_CUDABackend::sim_step
 * (src/CUDA/Backends/MD_CUDABackend.cu:567-619) driven by SimManager::run
 * (src/Managers/SimManager.cpp:158-180); steps are captured in a HIP graph
 * and the rebuild decision stays on the device. */
oxh_status oxh_batch_run(oxh_batch *b, long long n_steps);

/* Reorders particles of every system along a Hilbert curve (the
 * reference's _sort_particles, src/CUDA/Backends/MD_CUDABackend.cu:535-550,
 * with CUDA_sort_every).  Downloads keep the original particle order. */
oxh_status oxh_batch_hilbert_sort(oxh_batch *b);

/* Checkpoints: the batch's whole dynamical state -- FP64 positions,
 * velocities, angular momenta and quaternions, the positions of the last
 * list build, the forces of the current state, the per-particle RNG step
 * counters, the particle order, Bussi's internal kinetic energies, the step
 * counters and the force launch shape -- as one opaque byte buffer.
 * Restored into a batch created from the same systems (same descs), the
 * run continues bitwise as if it had not been interrupted: the lists are
 * rebuilt exactly from their build positions and the Philox streams
 * continue from the saved counters.  The reference's checkpoints
 * (checkpoint_every / checkpoint_file, src/Backends/SimBackend.cpp:250-266)
 * hold a configuration and the thermostat's RNG state; its CUDA backend
 * cannot resume a trajectory exactly (src/CUDA/Backends/CUDABaseBackend.cu:
 * 137-140).  Batches only, not slot pools. */
oxh_status oxh_batch_checkpoint_bytes(oxh_batch *b, size_t *bytes);
oxh_status oxh_batch_save(oxh_batch *b, void *buf, size_t bytes);
oxh_status oxh_batch_restore(oxh_batch *b, const void *buf, size_t bytes);

/* Re-sorts along the Hilbert curve every `steps` MD steps (0 = never),
 * starting before the next step; the reference's CUDA_sort_every
 * (src/CUDA/Backends/MD_CUDABackend.cu:575-577) counts list updates, this
 * counts steps so that the decision stays on the host between graph
 * chunks.  The interval is rounded up to whole chunks of the batch (64
 * steps, or a multiple of newtonian_steps with Bussi), and sorts fall on the
 * same steps however a run is split into oxh_batch_run calls (so that a
 * restored checkpoint continues bitwise). */
oxh_status oxh_batch_set_sort_interval(oxh_batch *b, long long steps);

/* Moves every strand whose centre of mass has left the box back into it by
 * whole box lengths (the reference's fix_diffusion, src/Backends/
 * SimBackend.cpp:788-882): positions and list positions move together;
 * undone if the potential energy changes by more than 1e-6 of itself.
 * n_moved (may be NULL) receives the number of strands moved. */
oxh_status oxh_batch_fix_diffusion(oxh_batch *b, int *n_moved);

/* Runs oxh_batch_fix_diffusion every `steps` MD steps of oxh_batch_run (0 =
 * never; the reference's fix_diffusion / fix_diffusion_every keys, default
 * on every 100000 steps, src/Managers/SimManager.cpp:92,151), rounded up to
 * whole chunks and applied at the same steps however a run is split into
 * calls.  Slot pools do not re-box. */
oxh_status oxh_batch_set_fix_diffusion(oxh_batch *b, long long steps);

/* Blocks until all queued work of the batch has finished. */
oxh_status oxh_batch_sync(oxh_batch *b);

/* State of system `sys` in its original particle order (any pointer may be
 * NULL).  Replaces MD_CUDABackend::apply_simulation_data_changes
 * (src/CUDA/Backends/MD_CUDABackend.cu:315-394). */
oxh_status oxh_batch_get_state(oxh_batch *b, int sys, double *pos, double *a1, double *a3, double *vel, double *L);

/* The current device order of system `sys`: order[k] (n entries) is the
 * original index of the particle held in slot k.  It is the identity until
 * a Hilbert sort, then the stable order of the particles' (cell) Hilbert
 * keys.  Replaces the host index maps the reference rebuilds after a sort
 * (_h_particles_hindex, src/CUDA/Backends/MD_CUDABackend.cu:323-327;
 * get_inverted_sorted_hindex, src/CUDA/CUDA_sort.cu:183-193). */
oxh_status oxh_batch_get_order(oxh_batch *b, int sys, int *order);

/* ---- external forces (the reference's external_forces = true with an
 * external_forces_file; ForceFactory::add_force, src/Forces/
 * ForceFactory.cpp:66-110, the CPU force classes under src/Forces/ and the
 * CUDA_trap family set by MD_CUDABackend::_set_external_forces,
 * src/CUDA/Backends/MD_CUDABackend.cu:528-533, CUDA_MD.cuh:97-554).
 * Types, by the reference's `type` key: */
#define OXH_EXT_STRING 1            /* "string": ConstantRateForce */
#define OXH_EXT_TRAP 2              /* "trap": MovingTrap */
#define OXH_EXT_MUTUAL_TRAP 3       /* "mutual_trap": MutualTrap */
#define OXH_EXT_CONSTANT_TRAP 4     /* "constant_trap": ConstantTrap */
#define OXH_EXT_REPULSION_PLANE 5   /* "repulsion_plane": RepulsionPlane */
#define OXH_EXT_SPHERE 6            /* "sphere": RepulsiveSphere */
#define OXH_EXT_LOWDIM_TRAP 7       /* "lowdim_trap": LowdimMovingTrap */
#define OXH_EXT_ATTRACTION_PLANE 8  /* "attraction_plane": AttractionPlane */
#define OXH_EXT_PLANE_MOVING 9      /* "repulsion_plane_moving": RepulsionPlaneMoving */
#define OXH_EXT_SPHERE_SMOOTH 10    /* "sphere_smooth": RepulsiveSphereSmooth */
#define OXH_EXT_SPHERE_MOVING 11    /* "repulsive_sphere_moving": RepulsiveSphereMoving */
#define OXH_EXT_KEPLER_POINSOT 12   /* "repulsive_kepler_poinsot": RepulsiveKeplerPoinsot */
#define OXH_EXT_LJ_WALL 13          /* "LJ_wall": LJWall */
#define OXH_EXT_TWIST 14            /* "twist": ConstantRateTorque */
#define OXH_EXT_CENTRAL 15          /* "generic_central_force": GenericCentralForce (force_type = gravity) */
#define OXH_EXT_LJ_CONE 16          /* "LJ_cone": LJCone */
#define OXH_EXT_ELLIPSOID 17        /* "ellipsoid": RepulsiveEllipsoid */
#define OXH_EXT_COM 18              /* "com": COMForce */
#define OXH_EXT_YUKAWA_SPHERE 19    /* "yukawa_sphere": YukawaSphere */

/* One force acting on one particle, with the reference's keys (vectors
 * called "normalised" are normalised here, as the reference's init does):
 *   string:          F0, rate, dir (normalised unless dir_as_centre, then
 *                    dir is the centre)
 *   trap:            stiff, rate, dir (normalised), pos0
 *   mutual_trap:     stiff, stiff_rate, r0, rate, ref_particle, PBC
 *   constant_trap:   stiff, r0, ref_particle, PBC
 *   repulsion_plane: stiff, position, end_position, v, dir (normalised)
 *   sphere:          stiff, r0, rate, r_ext, center
 *   lowdim_trap:     stiff, rate, pos0, dir (normalised), visibility
 *   attraction_plane: stiff, position, dir (normalised)
 *   repulsion_plane_moving: stiff, dir (normalised), the reference
 *                    particles ref_particle..ref_last (contiguous, as the
 *                    reference requires)
 *   sphere_smooth:   stiff, r0, r_ext, smooth, alpha, center
 *   repulsive_sphere_moving: stiff, r0, rate, r_ext, origin (in pos0),
 *                    target, steps
 *   repulsive_kepler_poinsot: stiff, rate, apex, base, base_radius, kappa,
 *                    center
 *   LJ_wall:         stiff, position, sigma, n, cutoff, dir (normalised)
 *   twist:           stiff, rate, base (in F0), axis (normalised), pos0,
 *                    center, mask
 *   generic_central_force: F0, inner_cut_off, outer_cut_off, center
 *   LJ_cone:         stiff, sigma, alpha, n, cutoff, dir (normalised),
 *                    pos0
 *   ellipsoid:       stiff, r_1, r_2, center
 *   com:             stiff, r0, rate, com_list (n_com), ref_list (n_ref):
 *                    every particle of com_list carries the record
 *   yukawa_sphere:   radius, epsilon (WCA_epsilon), sigma (WCA_sigma), n
 *                    (WCA_n), debye_length, debye_A, cutoff, center
 * A force the reference applies to a list of particles (particle = "0-9,
 * 12" or -1 for all) is one record per particle, in the file's order.
 * Keys are taken as the reference's init leaves them, defaults and key
 * aliases included (e.g. only_repulsive sets cutoff = 2^(1/n)). */
typedef struct oxh_ext_force {
	int type;
	int particle;        /* index in the system, upload order */
	int ref_particle;    /* mutual_trap, constant_trap, repulsion_plane_moving */
	int pbc;             /* mutual_trap, constant_trap: 1 = nearest image */
	int dir_as_centre;   /* string */
	double F0, rate, stiff, stiff_rate, r0;
	double position, end_position, v, r_ext;
	double dir[3];
	double pos0[3];
	int ref_last;        /* repulsion_plane_moving */
	int n;               /* LJ_wall, LJ_cone: exponent; yukawa_sphere: WCA_n */
	int visibility[3];   /* lowdim_trap */
	int n_com, n_ref;    /* com */
	const int *com_list; /* com: n_com indices in the system */
	const int *ref_list; /* com: n_ref indices in the system */
	long long steps;     /* repulsive_sphere_moving */
	double sigma, cutoff, alpha, smooth, epsilon, debye_length, debye_A, radius;
	double apex, base, base_radius, kappa, inner_cut_off, outer_cut_off;
	double center[3], axis[3], mask[3], target[3], r_1[3], r_2[3];
} oxh_ext_force;

/* Replaces the external forces of system `sys` (n = 0 removes them).
 * step0 is the reference's step counter at the system's next force
 * evaluation (the step read from the configuration, or 0 with
 * restart_step_counter): the forces of MD step k of later oxh_batch_run
 * calls use step0 + k, counted from the system's first step.  The forces
 * act from the current state on: F of the current state is recomputed.
 * Batches only (not slot pools). */
oxh_status oxh_batch_set_external_forces(oxh_batch *b, int sys, int n, const oxh_ext_force *forces, long long step0);

/* Periodic image counts of system `sys` (3 ints per particle, original
 * order): the absolute position of a particle is its position plus
 * image * box, as the reference's OrthogonalBox::get_abs_pos
 * (src/Boxes/OrthogonalBox.cpp:63-74); oxh_batch_fix_diffusion moves
 * positions by whole boxes and counts the moves here.  Set: e.g. from the
 * shift fields of a binary configuration (src/Backends/SimBackend.cpp:
 * 630-645). */
oxh_status oxh_batch_get_images(oxh_batch *b, int sys, int *img3);
oxh_status oxh_batch_set_images(oxh_batch *b, int sys, const int *img3);

/* Forces (lab) and torques (body frame) of the last force evaluation, and
 * each particle's share of the potential energy (Epart, N doubles). */
oxh_status oxh_batch_get_forces(oxh_batch *b, int sys, double *F, double *T, double *Epart);

/* ---- FFS order parameters.  Per-pair quantities of system `sys` at its
 * current state, FP64, for n pairs (2 ints each: particle indices in the
 * system's upload order), one value per pair into out.  They are what the
 * reference's order parameters are made of (src/Utilities/
 * OrderParameters.cpp, src/Observables/OrderParameterValues.cpp:55-66) and
 * what its CUDA FFS backend precomputes per pair (hb_op_precalc,
 * near_hb_op_precalc, dist_op_precalc, src/CUDA/Interactions/
 * CUDA_DNA.cuh:909-1069):
 *   OXH_PAIR_HB_ENERGY        hydrogen-bonding energy (0 for bonded
 *                             neighbours; a "bond" parameter counts pairs
 *                             below -0.1);
 *   OXH_PAIR_NEAR_HB          1 when the pair is complementary, its base-base
 *                             distance is in the HB range and at least 6 of
 *                             the 7 HB factors are active, else 0;
 *   OXH_PAIR_BASE_DISTANCE    |base(q) - base(p)| ("mindistance");
 *   OXH_PAIR_CENTRE_DISTANCE  |q - p| ("mindistance" with use_COM).
 * HB quantities take the nearest image; distances take the plain
 * difference for two particles of one strand and the nearest image
 * otherwise, as the reference's MinDistanceParameter. */
#define OXH_PAIR_HB_ENERGY 0
#define OXH_PAIR_NEAR_HB 1
#define OXH_PAIR_BASE_DISTANCE 2
#define OXH_PAIR_CENTRE_DISTANCE 3
oxh_status oxh_batch_pair_quantities(oxh_batch *b, int sys, int kind, int n, const int *pairs, double *out);

/* ---- barostat.  One Monte Carlo volume move of system `sys`, the
 * reference's MC barostat of MD runs (use_barostat, P, delta_L,
 * barostat_isotropic, barostat_molecular: src/Backends/MDBackend.cpp:43-60;
 * the moves of src/Backends/MCMoves/VolumeMove.cpp:49-122 and
 * MoleculeVolumeMove.cpp, applied by MD_CPUBackend, src/Backends/
 * MD_CPUBackend.cpp:200-205; the CUDA backend's MD_CUDABackend::
 * _apply_barostat, src/CUDA/Backends/MD_CUDABackend.cu:451-516).  Each box
 * side grows by delta_L (u - 1/2) (u = u4[0] for all three when
 * isotropic, else u4[0..2]); positions scale with the box (molecular:
 * every strand moves with its centre of mass); the lists are rebuilt and
 * the move is accepted when exp(-(dU + P dV - n T ln(V'/V)) / T) > u4[3],
 * with n = N (atomic) or the number of strands + 1 (molecular), the
 * CPU backend's counts.  A rejected move restores the box and positions.
 * u4 are uniform numbers in [0, 1) from the caller's generator (the
 * reference draws drand48).  accepted, dE (may be NULL): the decision and
 * dU.  Batches only; a system with external forces is refused. */
oxh_status oxh_batch_barostat_move(oxh_batch *b, int sys, double P, double delta_L, int isotropic, int molecular, const double *u4,
		int *accepted, double *dE);

/* Box sides of system `sys` (3 doubles). */
oxh_status oxh_batch_box(oxh_batch *b, int sys, double *box3);

/* Per-system FP64 sums: U (potential), U_hb, K_t, K_r (4 doubles per
 * system, n_sys entries).  Replaces the reference's CUDA energy print
 * (src/CUDA/Backends/MD_CUDABackend.cu:603-606, thrust transform_reduce). */
oxh_status oxh_batch_energies(oxh_batch *b, double *out4);

/* Potential energy of one system split by term, in the reference's order
 * (FENE, bonded excl, stacking, non-bonded excl, HB, cross stacking,
 * coaxial stacking, Debye-Hueckel), as BaseInteraction::
 * get_system_energy_split (src/Interactions/BaseInteraction.cpp:61-86). */
oxh_status oxh_batch_energy_split(oxh_batch *b, int sys, double *E8);

/* Stress tensor of system `sys` at its current state, components xx, yy,
 * zz, xy, xz, yz: virial6 = sum over interacting pairs of -r (x) F_p (r = q
 * - p centre separation, F_p the pair's force on p), kinetic6 = sum of
 * v (x) v.  (virial6 + kinetic6) / V is the reference's stress tensor and
 * its trace / 3 the pressure (DNAInteraction's _update_stress_tensor,
 * src/Interactions/BaseInteraction.cpp:159-200,285-352,
 * src/Observables/Pressure.cpp); the GPU counterpart is
 * CPU_stress_tensor (src/CUDA/Interactions/CUDABaseInteraction.cu:134-141).
 * Either pointer may be NULL. */
oxh_status oxh_batch_stress(oxh_batch *b, int sys, double *virial6, double *kinetic6);

/* Number of base pairs (pairs with E_HB < -0.1, the reference's HB_CUTOFF,
 * src/defs.h:62, src/Observables/HBList.cpp:141-164) of every system. */
oxh_status oxh_batch_hb_count(oxh_batch *b, int *out);

/* The current Verlet pair set of system `sys` as unordered pairs
 * (i < j, original indices), written to pairs (2 ints each) up to max
 * pairs; *n receives the total.  Exposes what the reference keeps in
 * CUDABaseList::d_matrix_neighs (src/CUDA/Lists/CUDABaseList.h:25-62):
 * the same set, except that a system whose far pairs are kept by their
 * backbone sites (oxh_batch_stats out[11]) leaves out the far pairs whose
 * Debye-Hueckel term is zero until the next rebuild. */
oxh_status oxh_batch_get_pairs(oxh_batch *b, int sys, int *pairs, long long max, long long *n);

/* List statistics of system `sys`: rebuild count and the last error bits
 * (1 broken FENE bond, 2 neighbour overflow, 4 cell overflow, 16 a
 * particle inside a yukawa_sphere's wall, where the reference throws). */
oxh_status oxh_batch_system_info(oxh_batch *b, int sys, long long *n_rebuilds, int *error);

/* Replaces the state (positions, orientations, velocities) of system `sys`
 * in place; topology, N and box stay.  Used to reuse a resident slot. */
oxh_status oxh_batch_set_state(oxh_batch *b, int sys, const double *pos, const double *a1, const double *a3, const double *vel,
		const double *L);

/* Neighbour-list, memory and launch-shape figures of the batch (11
 * doubles): out[0] N (a pool: its slots' total capacity), out[1]
 * max_neigh, out[2] mean neighbours per particle of the systems held,
 * out[3] device bytes held, out[4] threads per nucleotide and
 * out[5] block size of the close-pair force kernel, out[6] threads per
 * nucleotide (0 = inline) and out[7] block size of the far-pair kernel,
 * out[8] engine of oxh_batch_run (0: one kernel per phase, 1: one
 * workgroup per system for many steps per launch, used when every system
 * has at most 512 nucleotides and no cells, 2: one kernel per phase with
 * the close pairs in half-list form, each pair evaluated once -- the
 * reference's use_edge), out[9] number of systems, out[10] 1 when the
 * close kernel takes the packed form (HB, cross and coaxial stacking of a
 * workgroup's close pairs evaluated from work lists in LDS, one pair per
 * thread; same results bit for bit), else 0, out[11] number of systems
 * whose far pairs are kept by their backbone sites (far pairs then enter
 * the list only while their backbone sites are within the Debye-Hueckel
 * cut-off plus twice the verlet_skin, and a backbone site that moves more
 * than the skin triggers a rebuild too: the same forces from a shorter
 * list; OXH_FAR_BACK=0 keeps every far pair within r_verlet, as the
 * reference's list does).
 * Shapes are picked by timing at upload (the all-FP64 mode times the lanes
 * only, in blocks of 64 with the far pairs inline; pools keep 8 lanes);
 * OXH_FORCE_LANES / OXH_FAR_LANES fix them, OXH_PERSIST=0 disables the
 * workgroup engine, OXH_EDGE=0 / 1 forbids / forces the half-list form,
 * OXH_PACK=0 / 1 forbids / forces the packed form (never used with one
 * lane per nucleotide). */
oxh_status oxh_batch_stats(oxh_batch *b, double *out12);

/* Average duration in ms of n back-to-back launches of the force kernel on
 * the batch's stream, timed with HIP events (the reference times its steps
 * with CPU clocks around device syncs, src/Utilities/Timings.cpp:45-62). */
oxh_status oxh_batch_time_forces(oxh_batch *b, int n_launches, double *avg_ms);

/* ---- slot pools (served form).  The reference runs one system per
 * process (src/Utilities/ConfigInfo.h:47) and has no counterpart: a job
 * server keeps pools resident and admits jobs into free slots while the
 * other slots keep running (continuous batching).
 *
 * A pool is an oxh_batch of n_slots empty slots; slot s reserves
 * capacity[s] nucleotides and cell_budget[s] cells (NULL: capacity / 2
 * cells for slots above 1024 nucleotides, none below).  Pools whose slots
 * all hold at most 512 nucleotides run the workgroup engine.  Slots start
 * empty; oxh_batch_run advances every occupied slot by at most its
 * remaining steps, so jobs stop exactly at their targets.  A job's
 * trajectory depends only on the job: lanes per nucleotide are fixed per
 * pool, lists are rebuilt on each system's own skin trigger, and capacity
 * changes rebuild lists from the last build's positions.  All other
 * oxh_batch_* calls take a slot index as the system index. */
oxh_status oxh_pool_create(int device, int n_slots, const int *capacity, const int *cell_budget, oxh_batch **out);

/* Uploads job desc into empty slot `slot` (desc->N <= capacity), builds
 * its list and forces; it will run n_steps steps. */
oxh_status oxh_pool_admit(oxh_batch *pool, int slot, const oxh_system_desc *desc, long long n_steps);

/* Empties a slot (download results first). */
oxh_status oxh_pool_release(oxh_batch *pool, int slot);

/* Per slot: steps done, steps targeted, nucleotides (0 = empty) and error
 * bits (bit 0: a bonded pair exceeded the FENE range; that slot then stops
 * while the others continue).  Any pointer may be NULL. */
oxh_status oxh_pool_status(oxh_batch *pool, long long *steps_done, long long *steps_target, int *n_nucleotides, int *error);

/* Hilbert keys of n cells, each (x, y, z, bits) as 4 unsigned ints, on
 * the device: the same function the sort uses, for checking. */
oxh_status oxh_hilbert_keys(int device, int n, const unsigned int *xyzb, unsigned int *keys);

/* Philox4x32-10 on the device for n (counter[4], key[2]) inputs: the
 * thermostats' generator, for checking against published answers. */
oxh_status oxh_philox(int device, int n, const unsigned int *ctr4, const unsigned int *key2, unsigned int *out4);

/* Copies `bytes` between two device buffers with a plain float4 copy
 * kernel (k_copy_probe), reps times, and returns the achieved HBM GB/s
 * (read + write).  Calibrates HBM traffic counters (a copy of B bytes
 * fetches and writes B bytes) and measures the achievable bandwidth; no
 * counterpart in the reference. */
oxh_status oxh_copy_bench(int device, long long bytes, int reps, double *gbps);

#ifdef __cplusplus
}
#endif

#endif /* OXH_H */
