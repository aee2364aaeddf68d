"""ctypes binding of the C ABI in include/oxh.h (oxdna_amd/liboxh.so).

There is no fallback: if the HIP library is not built, or no GPU is
visible, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboxh.so")

OXH_THERMOSTAT = {"no": 0, "langevin": 1, "bussi": 2, "brownian": 3, "john": 3}
# interaction_type -> OXH_INTERACTION_*
OXH_INTERACTION = {"DNA2": 0, "DNA2_nomesh": 0, "RNA2": 1, "RNA": 2}

# every entry point declared in include/oxh.h
EXPORTS = [
    "oxh_system_desc_defaults", "oxh_last_error", "oxh_version", "oxh_device_count", "oxh_batch_create",
    "oxh_batch_destroy", "oxh_batch_update_lists", "oxh_batch_compute_forces", "oxh_batch_run",
    "oxh_batch_hilbert_sort", "oxh_batch_set_sort_interval", "oxh_batch_sync", "oxh_batch_get_state", "oxh_batch_get_forces", "oxh_batch_energies",
    "oxh_batch_energy_split", "oxh_batch_hb_count", "oxh_batch_get_pairs", "oxh_batch_system_info",
    "oxh_batch_set_state", "oxh_batch_stats", "oxh_batch_time_forces", "oxh_hilbert_keys", "oxh_philox",
    "oxh_pool_create", "oxh_pool_admit", "oxh_pool_release", "oxh_pool_status", "oxh_copy_bench", "oxh_batch_stress",
    "oxh_batch_fix_diffusion", "oxh_batch_set_fix_diffusion", "oxh_batch_checkpoint_bytes", "oxh_batch_save",
    "oxh_batch_restore", "oxh_batch_get_order", "oxh_batch_set_external_forces", "oxh_batch_get_images",
    "oxh_batch_set_images", "oxh_interaction_compute_forces", "oxh_interaction_compute_forces_noalloc",
    "oxh_batch_pair_quantities", "oxh_batch_barostat_move", "oxh_batch_box",
]

# per-pair quantities of FFS order parameters (include/oxh.h OXH_PAIR_*)
OXH_PAIR = {"hb_energy": 0, "near_hb": 1, "base_distance": 2, "centre_distance": 3}

# external force types (include/oxh.h OXH_EXT_*), by the reference's `type` key
OXH_EXT = {"string": 1, "trap": 2, "mutual_trap": 3, "constant_trap": 4, "repulsion_plane": 5, "sphere": 6,
           "lowdim_trap": 7, "attraction_plane": 8, "repulsion_plane_moving": 9, "sphere_smooth": 10,
           "repulsive_sphere_moving": 11, "repulsive_kepler_poinsot": 12, "LJ_wall": 13, "twist": 14,
           "generic_central_force": 15, "LJ_cone": 16, "ellipsoid": 17, "com": 18, "yukawa_sphere": 19}


class SystemDesc(C.Structure):
    _fields_ = [
        ("N", C.c_int),
        ("pos", C.c_void_p), ("a1", C.c_void_p), ("a3", C.c_void_p), ("vel", C.c_void_p), ("L", C.c_void_p),
        ("btype", C.c_void_p), ("n3", C.c_void_p), ("n5", C.c_void_p),
        ("box", C.c_double * 3),
        ("T", C.c_double), ("salt", C.c_double), ("dh_half_charged_ends", C.c_int),
        ("dh_lambda", C.c_double), ("dh_strength", C.c_double), ("dh_rhigh", C.c_double),
        ("hb_multiplier", C.c_double), ("max_backbone_force", C.c_double), ("max_backbone_force_far", C.c_double),
        ("seq_stck", C.c_void_p), ("seq_stck_fact_eps", C.c_double), ("seq_hydr_AT", C.c_double),
        ("seq_hydr_GC", C.c_double),
        ("verlet_skin", C.c_double), ("dt", C.c_double), ("thermostat", C.c_int), ("diff_coeff", C.c_double),
        ("gamma_trans", C.c_double), ("pt", C.c_double), ("newtonian_steps", C.c_int), ("bussi_tau", C.c_int),
        ("seed", C.c_ulonglong), ("refresh_vel", C.c_int), ("backend_precision", C.c_int),
        ("interaction", C.c_int),
    ]


class ExtForce(C.Structure):
    _fields_ = [
        ("type", C.c_int), ("particle", C.c_int), ("ref_particle", C.c_int), ("pbc", C.c_int),
        ("dir_as_centre", C.c_int),
        ("F0", C.c_double), ("rate", C.c_double), ("stiff", C.c_double), ("stiff_rate", C.c_double),
        ("r0", C.c_double), ("position", C.c_double), ("end_position", C.c_double), ("v", C.c_double),
        ("r_ext", C.c_double), ("dir", C.c_double * 3), ("pos0", C.c_double * 3),
        ("ref_last", C.c_int), ("n", C.c_int), ("visibility", C.c_int * 3), ("n_com", C.c_int), ("n_ref", C.c_int),
        ("com_list", C.c_void_p), ("ref_list", C.c_void_p), ("steps", C.c_longlong),
        ("sigma", C.c_double), ("cutoff", C.c_double), ("alpha", C.c_double), ("smooth", C.c_double),
        ("epsilon", C.c_double), ("debye_length", C.c_double), ("debye_A", C.c_double), ("radius", C.c_double),
        ("apex", C.c_double), ("base", C.c_double), ("base_radius", C.c_double), ("kappa", C.c_double),
        ("inner_cut_off", C.c_double), ("outer_cut_off", C.c_double),
        ("center", C.c_double * 3), ("axis", C.c_double * 3), ("mask", C.c_double * 3), ("target", C.c_double * 3),
        ("r_1", C.c_double * 3), ("r_2", C.c_double * 3),
    ]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m oxdna_amd.build`")
        L = C.CDLL(LIB_PATH)
        V, I, D, P, LL = C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_longlong
        L.oxh_last_error.restype = C.c_char_p
        L.oxh_system_desc_defaults.argtypes = [P]
        L.oxh_device_count.argtypes = [P]
        L.oxh_batch_create.argtypes = [I, I, P, P]
        L.oxh_batch_destroy.argtypes = [V]
        L.oxh_batch_update_lists.argtypes = [V, I]
        L.oxh_batch_compute_forces.argtypes = [V]
        L.oxh_interaction_compute_forces.argtypes = [V, P, P, P, P, P]
        L.oxh_interaction_compute_forces_noalloc.argtypes = [I, C.POINTER(SystemDesc), P, P, P, P, P]
        L.oxh_batch_run.argtypes = [V, LL]
        L.oxh_batch_hilbert_sort.argtypes = [V]
        L.oxh_batch_sync.argtypes = [V]
        L.oxh_batch_set_sort_interval.argtypes = [V, LL]
        L.oxh_batch_get_state.argtypes = [V, I, P, P, P, P, P]
        L.oxh_batch_get_forces.argtypes = [V, I, P, P, P]
        L.oxh_batch_get_order.argtypes = [V, I, P]
        L.oxh_batch_set_external_forces.argtypes = [V, I, I, P, LL]
        L.oxh_batch_get_images.argtypes = [V, I, P]
        L.oxh_batch_set_images.argtypes = [V, I, P]
        L.oxh_batch_energies.argtypes = [V, P]
        L.oxh_batch_pair_quantities.argtypes = [V, I, I, I, P, P]
        L.oxh_batch_barostat_move.argtypes = [V, I, D, D, I, I, P, P, P]
        L.oxh_batch_box.argtypes = [V, I, P]
        L.oxh_batch_energy_split.argtypes = [V, I, P]
        L.oxh_batch_hb_count.argtypes = [V, P]
        L.oxh_batch_get_pairs.argtypes = [V, I, P, LL, P]
        L.oxh_batch_system_info.argtypes = [V, I, P, P]
        L.oxh_batch_set_state.argtypes = [V, I, P, P, P, P, P]
        L.oxh_batch_stats.argtypes = [V, P]
        L.oxh_batch_time_forces.argtypes = [V, I, P]
        L.oxh_hilbert_keys.argtypes = [I, I, P, P]
        L.oxh_philox.argtypes = [I, I, P, P, P]
        L.oxh_pool_create.argtypes = [I, I, P, P, C.POINTER(C.c_void_p)]
        L.oxh_pool_admit.argtypes = [V, I, C.POINTER(SystemDesc), LL]
        L.oxh_pool_release.argtypes = [V, I]
        L.oxh_pool_status.argtypes = [V, P, P, P, P]
        L.oxh_copy_bench.argtypes = [I, LL, I, P]
        L.oxh_batch_stress.argtypes = [V, I, P, P]
        L.oxh_batch_fix_diffusion.argtypes = [V, P]
        L.oxh_batch_set_fix_diffusion.argtypes = [V, LL]
        L.oxh_batch_checkpoint_bytes.argtypes = [V, P]
        L.oxh_batch_save.argtypes = [V, P, C.c_size_t]
        L.oxh_batch_restore.argtypes = [V, P, C.c_size_t]
        for name in EXPORTS:
            f = getattr(L, name)
            if name not in ("oxh_last_error", "oxh_system_desc_defaults"):
                f.restype = C.c_int
        L.oxh_system_desc_defaults.restype = None
        _LIB = L
    return _LIB


class OxhError(RuntimeError):
    pass


def check(status: int, what: str = ""):
    if status != 0:
        msg = lib().oxh_last_error().decode(errors="replace")
        raise OxhError(f"{what}: oxh status {status}: {msg}")


def device_count() -> int:
    n = C.c_int(0)
    check(lib().oxh_device_count(C.byref(n)), "oxh_device_count")
    return n.value


def ptr(a):
    return a.ctypes.data if a is not None else None
