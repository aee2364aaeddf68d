"""Python host side of the HIP backend: systems resident on the GPU.

`HIPBatch` owns one oxh_batch (include/oxh.h) holding one or more
independent systems.  It mirrors what the reference's MD backend does with
its device state (src/CUDA/Backends/MD_CUDABackend.cu): upload, list
updates, force evaluation, MD steps and downloads, but for many systems at
once.  The plain path is a batch of one system.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import SystemDesc, check, lib, ptr
from .systems import System

ENERGY_TERMS = ("FENE", "BONDED_EXCLUDED_VOLUME", "STACKING", "NONBONDED_EXCLUDED_VOLUME", "HYDROGEN_BONDING",
                "CROSS_STACKING", "COAXIAL_STACKING", "DEBYE_HUCKEL")


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def quat_from_frame(a1, a3) -> np.ndarray:
    """(N, 4) unit quaternions (x, y, z, w) of the rotations whose matrices
    have the columns a1, a2 = a3 x a1, a3 (the reference's p->orientation),
    the layout oxh_interaction_compute_forces reads.  Shepperd's method:
    the largest of 1 + trace and 1 + 2 M_kk - trace picks the component
    computed from a square root."""
    a1 = np.asarray(a1, np.float64)
    a3 = np.asarray(a3, np.float64)
    a2 = np.cross(a3, a1)
    M = np.stack([a1, a2, a3], axis=2)   # M[:, row, col], columns a1, a2, a3
    tr = M[:, 0, 0] + M[:, 1, 1] + M[:, 2, 2]
    cand = np.stack([tr, M[:, 0, 0], M[:, 1, 1], M[:, 2, 2]], axis=1)
    pick = np.argmax(cand, axis=1)
    q = np.zeros((len(a1), 4))
    for k in range(4):
        sel = pick == k
        m = M[sel]
        if k == 0:
            s = 2.0 * np.sqrt(1.0 + tr[sel])
            q[sel] = np.stack([(m[:, 2, 1] - m[:, 1, 2]) / s, (m[:, 0, 2] - m[:, 2, 0]) / s,
                               (m[:, 1, 0] - m[:, 0, 1]) / s, 0.25 * s], axis=1)
        else:
            i = k - 1
            j, l = (i + 1) % 3, (i + 2) % 3
            s = 2.0 * np.sqrt(1.0 + m[:, i, i] - m[:, j, j] - m[:, l, l])
            comp = np.zeros((len(m), 4))
            comp[:, i] = 0.25 * s
            comp[:, j] = (m[:, j, i] + m[:, i, j]) / s
            comp[:, l] = (m[:, l, i] + m[:, i, l]) / s
            comp[:, 3] = (m[:, l, j] - m[:, j, l]) / s
            q[sel] = comp
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def make_desc(s: System, keep: list, **opts) -> SystemDesc:
    """oxh_system_desc of System s; option names are the reference's
    input keys (verlet_skin, dt, thermostat, diff_coeff, newtonian_steps,
    bussi_tau, salt_concentration, dh_lambda, ...)."""
    d = SystemDesc()
    lib().oxh_system_desc_defaults(C.byref(d))
    arrs = dict(pos=_c(s.pos, np.float64), a1=_c(s.a1, np.float64), a3=_c(s.a3, np.float64),
                btype=_c(s.btype, np.int32), n3=_c(s.n3, np.int32), n5=_c(s.n5, np.int32))
    if s.vel is not None and not opts.get("refresh_vel", False):
        arrs["vel"] = _c(s.vel, np.float64)
    if s.L is not None and not opts.get("refresh_vel", False):
        arrs["L"] = _c(s.L, np.float64)
    keep.append(arrs)
    d.N = s.N
    for k, a in arrs.items():
        setattr(d, k, a.ctypes.data)
    for k in range(3):
        d.box[k] = float(s.box[k])
    d.T = float(opts.get("T", s.T))
    d.salt = float(opts.get("salt_concentration", s.salt))
    d.dh_half_charged_ends = int(bool(opts.get("dh_half_charged_ends", True)))
    d.dh_lambda = float(opts.get("dh_lambda", -1.0))
    d.dh_strength = float(opts.get("dh_strength", -1.0))
    d.dh_rhigh = float(opts.get("debye_huckel_rhigh", -1.0))
    d.hb_multiplier = float(opts.get("hb_multiplier", 1.0))
    d.max_backbone_force = float(opts.get("max_backbone_force", -1.0))
    d.max_backbone_force_far = float(opts.get("max_backbone_force_far", -1.0))
    d.verlet_skin = float(opts.get("verlet_skin", 0.05))
    d.dt = float(opts.get("dt", 0.003))
    th = opts.get("thermostat", "no")
    d.thermostat = _lib.OXH_THERMOSTAT[th] if isinstance(th, str) else int(th)
    d.diff_coeff = float(opts.get("diff_coeff", -1.0))
    d.gamma_trans = float(opts.get("gamma_trans", -1.0))
    d.pt = float(opts.get("pt", -1.0))
    d.newtonian_steps = int(opts.get("newtonian_steps", 1))
    d.bussi_tau = int(opts.get("bussi_tau", 1))
    d.seed = int(opts.get("seed", 1)) & 0xFFFFFFFFFFFFFFFF
    d.refresh_vel = int(bool(opts.get("refresh_vel", False)))
    prec = opts.get("backend_precision", "mixed")
    if prec not in ("mixed", "double"):
        raise ValueError(f"backend_precision {prec!r}: use 'mixed' or 'double'")
    d.backend_precision = 1 if prec == "double" else 0
    it = opts.get("interaction_type", "DNA2")
    if it not in _lib.OXH_INTERACTION:
        raise ValueError(f"interaction_type {it!r}: use one of {sorted(_lib.OXH_INTERACTION)}")
    d.interaction = _lib.OXH_INTERACTION[it]
    if "seq_stck" in opts:
        st = _c(opts["seq_stck"], np.float64).ravel()
        keep.append(st)
        d.seq_stck = st.ctypes.data
        d.seq_stck_fact_eps = float(opts["seq_stck_fact_eps"])
        d.seq_hydr_AT = float(opts["seq_hydr_AT"])
        d.seq_hydr_GC = float(opts["seq_hydr_GC"])
    return d


class HIPBatch:
    """Independent systems resident on one GPU."""

    def __init__(self, systems, device: int = 0, options=None, **common):
        if isinstance(systems, System):
            systems = [systems]
        self.systems = list(systems)
        if options is None:
            options = [{} for _ in self.systems]
        keep: list = []
        descs = (SystemDesc * len(self.systems))()
        for k, s in enumerate(self.systems):
            o = dict(common)
            o.update(options[k])
            descs[k] = make_desc(s, keep, **o)
        h = C.c_void_p()
        check(lib().oxh_batch_create(device, len(self.systems), descs, C.byref(h)), "oxh_batch_create")
        self.h = h
        self.device = device

    # -- lifecycle
    def close(self):
        if getattr(self, "h", None):
            lib().oxh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def nsys(self):
        return len(self.systems)

    # -- operations
    def run(self, steps: int):
        check(lib().oxh_batch_run(self.h, int(steps)), "oxh_batch_run")

    def compute_forces(self):
        check(lib().oxh_batch_compute_forces(self.h), "oxh_batch_compute_forces")

    def update_lists(self, force: bool = True):
        check(lib().oxh_batch_update_lists(self.h, int(force)), "oxh_batch_update_lists")

    def hilbert_sort(self):
        check(lib().oxh_batch_hilbert_sort(self.h), "oxh_batch_hilbert_sort")

    def fix_diffusion(self) -> int:
        """Re-boxes strands whose centre of mass left the box; returns how
        many moved."""
        n = C.c_int(0)
        check(lib().oxh_batch_fix_diffusion(self.h, C.byref(n)), "oxh_batch_fix_diffusion")
        return n.value

    def set_fix_diffusion(self, every: int):
        """Re-box strands every `every` steps of run() (0 = never)."""
        check(lib().oxh_batch_set_fix_diffusion(self.h, int(every)), "oxh_batch_set_fix_diffusion")

    def checkpoint(self) -> bytes:
        """The whole dynamical state as bytes (oxh_batch_save)."""
        n = C.c_size_t(0)
        check(lib().oxh_batch_checkpoint_bytes(self.h, C.byref(n)), "oxh_batch_checkpoint_bytes")
        buf = C.create_string_buffer(n.value)
        check(lib().oxh_batch_save(self.h, buf, n.value), "oxh_batch_save")
        return buf.raw

    def restore(self, data: bytes):
        """Continues from a checkpoint of a batch built from the same
        systems (oxh_batch_restore)."""
        buf = C.create_string_buffer(data, len(data))
        check(lib().oxh_batch_restore(self.h, buf, len(data)), "oxh_batch_restore")

    def set_sort_interval(self, steps: int):
        """Hilbert re-sort every `steps` MD steps (0 = never)."""
        check(lib().oxh_batch_set_sort_interval(self.h, int(steps)), "oxh_batch_set_sort_interval")

    def sync(self):
        check(lib().oxh_batch_sync(self.h), "oxh_batch_sync")

    # -- downloads
    def state(self, sys: int = 0) -> dict:
        n = self.systems[sys].N
        out = {k: np.zeros((n, 3)) for k in ("pos", "a1", "a3", "vel", "L")}
        check(lib().oxh_batch_get_state(self.h, sys, *(ptr(out[k]) for k in ("pos", "a1", "a3", "vel", "L"))),
              "oxh_batch_get_state")
        return out

    def set_external_forces(self, forces, sys: int = 0, step0: int = 0):
        """External forces of system `sys`: a list of dicts whose keys are
        the fields of the C record (type, particle, F0, rate, stiff, dir,
        center, ...; see oxdna_amd.io.read_external_forces), one record per
        particle, in the file's order.  step0 is the reference's step
        counter at the next force evaluation."""
        arr = (_lib.ExtForce * max(1, len(forces)))()
        keep = []   # com lists stay alive through the call
        scalars = (("F0", 0.0), ("rate", 0.0), ("stiff", 0.0), ("stiff_rate", 0.0), ("r0", 0.0), ("position", 0.0),
                   ("end_position", 1e6), ("v", 0.0), ("r_ext", 1e10), ("sigma", 1.0), ("cutoff", 1e6), ("alpha", 0.0),
                   ("smooth", 0.0), ("epsilon", 1.0), ("debye_length", 0.0), ("debye_A", 0.0), ("radius", 0.0),
                   ("apex", 0.0), ("base", 0.0), ("base_radius", 0.0), ("kappa", 0.0), ("inner_cut_off", 0.0),
                   ("outer_cut_off", 0.0))
        for k, f in enumerate(forces):
            e = arr[k]
            e.type = _lib.OXH_EXT[f["type"]]
            e.particle = int(f["particle"])
            e.ref_particle = int(f.get("ref_particle", -1))
            e.ref_last = int(f.get("ref_last", -1))
            e.pbc = int(bool(f.get("PBC", False)))
            e.dir_as_centre = int(bool(f.get("dir_as_centre", False)))
            e.n = int(f.get("n", 6))
            e.steps = int(f.get("steps", 0))
            for key, default in scalars:
                setattr(e, key, float(f.get(key, default)))
            for key in ("dir", "pos0", "center", "axis", "mask", "target", "r_1", "r_2"):
                v = f.get(key, (0.0, 0.0, 0.0))
                for c in range(3):
                    getattr(e, key)[c] = float(v[c])
            for c, v in enumerate(f.get("visibility", (0, 0, 0))):
                e.visibility[c] = int(v)
            if f["type"] == "com":
                com, ref = _c(f["com_list"], np.int32), _c(f["ref_list"], np.int32)
                keep += [com, ref]
                e.n_com, e.n_ref = len(com), len(ref)
                e.com_list, e.ref_list = ptr(com), ptr(ref)
        check(lib().oxh_batch_set_external_forces(self.h, sys, len(forces), C.byref(arr), int(step0)),
              "oxh_batch_set_external_forces")

    def images(self, sys: int = 0) -> np.ndarray:
        """(N, 3) periodic image counts: absolute position = pos + img * box."""
        out = np.zeros((self.systems[sys].N, 3), np.int32)
        check(lib().oxh_batch_get_images(self.h, sys, ptr(out)), "oxh_batch_get_images")
        return out

    def set_images(self, img, sys: int = 0):
        a = _c(img, np.int32)
        check(lib().oxh_batch_set_images(self.h, sys, ptr(a)), "oxh_batch_set_images")

    def order(self, sys: int = 0) -> np.ndarray:
        """Original index of the particle in each device slot of system
        `sys` (the identity until a Hilbert sort)."""
        out = np.zeros(self.systems[sys].N, np.int32)
        check(lib().oxh_batch_get_order(self.h, sys, ptr(out)), "oxh_batch_get_order")
        return out

    def interaction_forces(self, d_pos: int, d_quat: int, d_forces: int, d_torques: int, stream: int = 0):
        """Adds the pair forces and torques of the state in the caller's
        device buffers to the caller's device buffers
        (oxh_interaction_compute_forces): d_pos and d_quat hold 4 doubles
        per particle (see quat_from_frame), d_forces and d_torques 4 floats
        (mixed) or 4 doubles (double) per particle, in upload order.  All
        arguments are device addresses (ints); stream is a hipStream_t
        handle or 0 for the batch's own."""
        check(lib().oxh_interaction_compute_forces(self.h, d_pos, d_quat, d_forces, d_torques, stream or None),
              "oxh_interaction_compute_forces")

    def forces(self, sys: int = 0):
        n = self.systems[sys].N
        F = np.zeros((n, 3))
        T = np.zeros((n, 3))
        E = np.zeros(n)
        check(lib().oxh_batch_get_forces(self.h, sys, ptr(F), ptr(T), ptr(E)), "oxh_batch_get_forces")
        return F, T, E

    def energies(self) -> np.ndarray:
        """(nsys, 4): U, U_hb, K_t, K_r, FP64 sums."""
        out = np.zeros((self.nsys, 4))
        check(lib().oxh_batch_energies(self.h, ptr(out)), "oxh_batch_energies")
        return out

    def pair_quantities(self, kind: str, pairs, sys: int = 0) -> np.ndarray:
        """Per-pair FFS order-parameter quantities of system `sys`
        (oxh_batch_pair_quantities): kind 'hb_energy', 'near_hb',
        'base_distance' or 'centre_distance'; pairs (n, 2) particle indices
        in the system's upload order."""
        pr = _c(np.asarray(pairs, dtype=np.int32).reshape(-1, 2), np.int32)
        out = np.zeros(len(pr))
        check(lib().oxh_batch_pair_quantities(self.h, sys, _lib.OXH_PAIR[kind], len(pr), ptr(pr), ptr(out)),
              "oxh_batch_pair_quantities")
        return out

    def barostat_move(self, P: float, delta_L: float, u4, isotropic: bool = True, molecular: bool = False,
                      sys: int = 0):
        """One Monte Carlo volume move (oxh_batch_barostat_move) with the
        uniform numbers u4 (4 values in [0, 1)); returns (accepted, dU)."""
        u = _c(np.asarray(u4, dtype=np.float64).reshape(4), np.float64)
        acc = C.c_int(0)
        dE = C.c_double(0.0)
        check(lib().oxh_batch_barostat_move(self.h, sys, float(P), float(delta_L), int(bool(isotropic)), int(bool(molecular)),
                                            ptr(u), C.byref(acc), C.byref(dE)), "oxh_batch_barostat_move")
        return bool(acc.value), dE.value

    def box(self, sys: int = 0) -> np.ndarray:
        out = np.zeros(3)
        check(lib().oxh_batch_box(self.h, sys, ptr(out)), "oxh_batch_box")
        return out

    def energy_split(self, sys: int = 0) -> np.ndarray:
        out = np.zeros(8)
        check(lib().oxh_batch_energy_split(self.h, sys, ptr(out)), "oxh_batch_energy_split")
        return out

    def stress(self, sys: int = 0):
        """(virial, kinetic): 6 components each (xx, yy, zz, xy, xz, yz),
        not divided by the volume."""
        w = np.zeros(6)
        k = np.zeros(6)
        check(lib().oxh_batch_stress(self.h, sys, ptr(w), ptr(k)), "oxh_batch_stress")
        return w, k

    def hb_count(self) -> np.ndarray:
        out = np.zeros(self.nsys, np.int32)
        check(lib().oxh_batch_hb_count(self.h, ptr(out)), "oxh_batch_hb_count")
        return out

    def pairs(self, sys: int = 0) -> np.ndarray:
        n = C.c_longlong(0)
        cap = max(1024, self.systems[sys].N * 64)
        while True:
            buf = np.zeros((cap, 2), np.int32)
            check(lib().oxh_batch_get_pairs(self.h, sys, ptr(buf), cap, C.byref(n)), "oxh_batch_get_pairs")
            if n.value <= cap:
                return buf[: n.value]
            cap = n.value

    def info(self, sys: int = 0):
        nr = C.c_longlong(0)
        err = C.c_int(0)
        check(lib().oxh_batch_system_info(self.h, sys, C.byref(nr), C.byref(err)), "oxh_batch_system_info")
        return nr.value, err.value

    def set_state(self, sys: int, pos=None, a1=None, a3=None, vel=None, L=None):
        arrs = [None if a is None else _c(a, np.float64) for a in (pos, a1, a3, vel, L)]
        check(lib().oxh_batch_set_state(self.h, sys, *(ptr(a) for a in arrs)), "oxh_batch_set_state")

    def time_forces(self, n: int = 20) -> float:
        """Average force-kernel launch duration (ms), HIP events on the
        batch's stream."""
        ms = C.c_double(0.0)
        check(lib().oxh_batch_time_forces(self.h, int(n), C.byref(ms)), "oxh_batch_time_forces")
        return ms.value

    def stats(self):
        out = np.zeros(12)
        check(lib().oxh_batch_stats(self.h, ptr(out)), "oxh_batch_stats")
        return dict(N=int(out[0]), max_neigh=int(out[1]), mean_neigh=out[2], device_bytes=int(out[3]),
                    lanes=int(out[4]), block=int(out[5]), far_lanes=int(out[6]), far_block=int(out[7]),
                    engine={0: "grid", 1: "workgroup", 2: "grid-edge"}[int(out[8])], n_systems=int(out[9]),
                    pack=bool(out[10]), far_back=int(out[11]))


class HIPPool(HIPBatch):
    """A slot pool (oxh_pool_*): slots of fixed capacity that jobs are
    admitted into and released from while the other slots keep running.
    Pools of slots of at most 512 nucleotides run the workgroup engine."""

    def __init__(self, capacity, device: int = 0, cell_budget=None):
        cap = np.ascontiguousarray(capacity, np.int32)
        cb = None if cell_budget is None else np.ascontiguousarray(cell_budget, np.int32)
        h = C.c_void_p()
        check(lib().oxh_pool_create(device, len(cap), ptr(cap), None if cb is None else ptr(cb), C.byref(h)),
              "oxh_pool_create")
        self.h = h
        self.device = device
        self.capacity = cap
        self.systems = [None] * len(cap)

    @property
    def nsys(self):
        return len(self.capacity)

    def admit(self, slot: int, system: System, steps: int, **opts):
        keep: list = []
        d = make_desc(system, keep, **opts)
        check(lib().oxh_pool_admit(self.h, int(slot), C.byref(d), int(steps)), "oxh_pool_admit")
        self.systems[slot] = system

    def release(self, slot: int):
        check(lib().oxh_pool_release(self.h, int(slot)), "oxh_pool_release")
        self.systems[slot] = None

    def status(self):
        n = self.nsys
        done = np.zeros(n, np.int64)
        target = np.zeros(n, np.int64)
        nn = np.zeros(n, np.int32)
        err = np.zeros(n, np.int32)
        check(lib().oxh_pool_status(self.h, ptr(done), ptr(target), ptr(nn), ptr(err)), "oxh_pool_status")
        return dict(done=done, target=target, n=nn, error=err)


def copy_bench(nbytes: int = 1 << 30, reps: int = 5, device: int = 0) -> float:
    """Achieved HBM GB/s (read + write) of a plain device copy of nbytes."""
    out = C.c_double(0.0)
    check(lib().oxh_copy_bench(device, int(nbytes), int(reps), C.byref(out)), "oxh_copy_bench")
    return out.value


def hilbert_keys(xyz: np.ndarray, bits: int, device: int = 0) -> np.ndarray:
    xyzb = np.zeros((len(xyz), 4), np.uint32)
    xyzb[:, :3] = xyz
    xyzb[:, 3] = bits
    out = np.zeros(len(xyz), np.uint32)
    check(lib().oxh_hilbert_keys(device, len(xyz), ptr(xyzb), ptr(out)), "oxh_hilbert_keys")
    return out


def philox(ctr: np.ndarray, key: np.ndarray, device: int = 0) -> np.ndarray:
    ctr = np.ascontiguousarray(ctr, np.uint32).reshape(-1, 4)
    key = np.ascontiguousarray(key, np.uint32).reshape(-1, 2)
    out = np.zeros_like(ctr)
    check(lib().oxh_philox(device, len(ctr), ptr(ctr), ptr(key), ptr(out)), "oxh_philox")
    return out


def interaction_forces_noalloc(system: System, d_pos: int, d_quat: int, d_forces: int, d_torques: int, stream: int = 0,
                               device: int = 0, **opts):
    """oxh_interaction_compute_forces_noalloc: adds the pair forces and
    torques of one system's state (caller device buffers, layouts as in
    HIPBatch.interaction_forces) to the caller's device buffers, with no
    device allocation (the library's fixed scratch array; at most
    OXH_NOALLOC_MAX_N = 4096 nucleotides).  Topology and parameters come
    from `system` and the options, as for HIPBatch; stream 0 is the null
    stream."""
    keep: list = []
    d = make_desc(system, keep, **opts)
    check(lib().oxh_interaction_compute_forces_noalloc(device, C.byref(d), d_pos, d_quat, d_forces, d_torques, stream or None),
          "oxh_interaction_compute_forces_noalloc")
