"""ctypes binding of the FP64 CPU oracle (oracle/liboxoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, as the checker.  The product never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

DP = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
IP = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboxoracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboxoracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        L.oxo_params_size.restype = C.c_int
        L.oxo_params_init.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_double]
        L.oxo_params_set_seqdep.argtypes = [C.c_void_p, DP, C.c_double, C.c_double, C.c_double]
        L.oxo_params_rcut.argtypes = [C.c_void_p]
        L.oxo_params_rcut.restype = C.c_double
        L.oxo_params_dh.argtypes = [C.c_void_p, DP]
        L.oxo_forces.argtypes = [C.c_void_p, C.c_int, DP, DP, DP, DP, IP, IP, IP, DP, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p]
        L.oxo_forces.restype = C.c_double
        L.oxo_virial.argtypes = [C.c_void_p, C.c_int, DP, DP, DP, DP, IP, IP, IP, DP, C.c_void_p]
        L.oxo_virial.restype = C.c_double
        L.oxo_count_hb.argtypes = [C.c_void_p, C.c_int, DP, DP, DP, DP, IP, IP, IP, DP]
        L.oxo_count_hb.restype = C.c_int
        L.oxo_verlet_pairs.argtypes = [C.c_int, DP, IP, IP, DP, C.c_double, C.c_void_p, C.c_longlong]
        L.oxo_verlet_pairs.restype = C.c_longlong
        L.oxo_hilbert_key.argtypes = [C.c_uint, C.c_uint, C.c_uint, C.c_int]
        L.oxo_hilbert_key.restype = C.c_uint
        L.oxo_philox4x32_10.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.oxo_md_step.argtypes = [C.c_void_p, C.c_int, DP, DP, DP, DP, DP, DP, IP, IP, IP, DP, C.c_double, DP, DP]
        L.oxo_md_step.restype = C.c_double
        _LIB = L
    return _LIB


class Params:
    """oxDNA2 model constants of one system (oxo_params)."""

    def __init__(self, T, salt=0.5, half_charged_ends=True, dh_lambda=-1.0, dh_strength=-1.0, dh_rhigh=-1.0,
                 hb_multiplier=1.0, max_backbone_force=-1.0, max_backbone_force_far=-1.0):
        L = lib()
        self.buf = C.create_string_buffer(L.oxo_params_size())
        L.oxo_params_init(self.buf, T, salt, int(half_charged_ends), dh_lambda, dh_strength, dh_rhigh, hb_multiplier,
                          max_backbone_force, max_backbone_force_far)

    @property
    def rcut(self):
        return lib().oxo_params_rcut(self.buf)

    def dh(self):
        out = np.zeros(4)
        lib().oxo_params_dh(self.buf, out)
        return dict(RHIGH=out[0], RC=out[1], B=out[2], minus_kappa=out[3])

    def set_seqdep(self, stck16, stck_fact_eps, hydr_AT, hydr_GC):
        lib().oxo_params_set_seqdep(self.buf, np.ascontiguousarray(stck16, np.float64).ravel(), stck_fact_eps,
                                    hydr_AT, hydr_GC)


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def forces(params: Params, s):
    """Returns dict(E8, U, F, T (body frame), Epart) for a System-like object."""
    N = s.N
    a2 = np.cross(s.a3, s.a1)
    E8 = np.zeros(8)
    F = np.zeros((N, 3))
    T = np.zeros((N, 3))
    Ep = np.zeros(N)
    U = lib().oxo_forces(params.buf, N, _c(s.pos), _c(s.a1), _c(a2), _c(s.a3), _c(s.n3, np.int32),
                         _c(s.n5, np.int32), _c(s.btype, np.int32), _c(s.box), E8.ctypes.data, F.ctypes.data,
                         T.ctypes.data, Ep.ctypes.data)
    return dict(E8=E8, U=U, F=F, T=T, Epart=Ep)


def virial(params: Params, s) -> np.ndarray:
    """Pair virial of the standard stress tensor: sum over interacting pairs
    of -r (x) F_p (xx, yy, zz, xy, xz, yz), not divided by the volume."""
    a2 = np.cross(s.a3, s.a1)
    W = np.zeros(6)
    lib().oxo_virial(params.buf, s.N, _c(s.pos), _c(s.a1), _c(a2), _c(s.a3), _c(s.n3, np.int32), _c(s.n5, np.int32),
                     _c(s.btype, np.int32), _c(s.box), W.ctypes.data)
    return W


def count_hb(params: Params, s) -> int:
    a2 = np.cross(s.a3, s.a1)
    return lib().oxo_count_hb(params.buf, s.N, _c(s.pos), _c(s.a1), _c(a2), _c(s.a3), _c(s.n3, np.int32),
                              _c(s.n5, np.int32), _c(s.btype, np.int32), _c(s.box))


def verlet_pairs(pos, n3, n5, box, sqr_rverlet, max_pairs=None):
    """Unordered pairs (i < j), sorted, of the FP64 Verlet predicate."""
    N = len(pos)
    if max_pairs is None:
        max_pairs = max(1024, N * 64)
    while True:
        out = np.zeros((max_pairs, 2), np.int32)
        n = lib().oxo_verlet_pairs(N, _c(pos), _c(n3, np.int32), _c(n5, np.int32), _c(box), sqr_rverlet,
                                   out.ctypes.data, max_pairs)
        if n <= max_pairs:
            out = out[:n]
            order = np.lexsort((out[:, 1], out[:, 0]))
            return out[order]
        max_pairs = int(n)


def hilbert_key(x, y, z, bits):
    return lib().oxo_hilbert_key(int(x), int(y), int(z), int(bits))


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().oxo_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def md_step(params: Params, s, dt, F, T, vel, L):
    """One FP64 velocity-Verlet step in place on s.pos/s.a1/s.a3, vel, L;
    F and T (body) hold the current forces and are updated.  Returns U."""
    a2 = np.ascontiguousarray(np.cross(s.a3, s.a1))
    pos, a1, a3 = _c(s.pos).copy(), _c(s.a1).copy(), _c(s.a3).copy()
    U = lib().oxo_md_step(params.buf, s.N, pos, a1, a2, a3, vel, L, _c(s.n3, np.int32), _c(s.n5, np.int32),
                          _c(s.btype, np.int32), _c(s.box), dt, F, T)
    s.pos[:], s.a1[:], s.a3[:] = pos, a1, a3
    return U
