"""The accumulating GPU interaction surface (oxh_interaction_compute_forces)
with caller-owned device buffers, as the reference's
CUDABaseInteraction::compute_forces (src/CUDA/Interactions/
CUDABaseInteraction.h:60): the pair forces and torques of the caller's
state are ADDED to what the caller's F and T hold, as after
MD_CUDABackend::_set_external_forces (src/CUDA/Backends/
MD_CUDABackend.cu:518-533).

Checked through the C ABI with device buffers from the HIP runtime
(tests/_hipmem.py) as the caller's memory: the additions equal pre-set values + the batch's own forces bit for
bit (FP32 additions in the mixed form, FP64 in the double form), the
forces match the FP64 oracle at the force tests' bars, the energy shares
in .w sum to the potential energy, a state moved by more than the Verlet
skin gets new lists (the oracle agrees after the move), and the caller's
own stream orders the work.
"""
import numpy as np
import pytest

from _common import load_golden, oracle_params, vec_rel_err
from _hipmem import DeviceArray, Stream

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from oxdna_amd import _lib, backend
    assert _lib.device_count() >= 1, "no GPU visible"
    return backend


def _state4(hip, s):
    pos4 = np.zeros((s.N, 4))
    pos4[:, :3] = s.pos
    return DeviceArray(pos4), DeviceArray(hip.quat_from_frame(s.a1, s.a3))


@pytest.mark.parametrize("precision", ["mixed", "double"])
@pytest.mark.parametrize("name", ["hairpin_md_50C", "dense_md_excl_30C", "bundle_md_30C"])
def test_accumulates_onto_preset_forces(hip, name, precision):
    from oracle import pyoracle as po
    s, opts, ref = load_golden(name)
    ftype = np.float32 if precision == "mixed" else np.float64
    tol = 1e-5 if precision == "mixed" else 1e-9
    rng = np.random.default_rng(5)
    ext_F = rng.normal(size=(s.N, 4))
    ext_T = rng.normal(size=(s.N, 4))
    with hip.HIPBatch(s, backend_precision=precision, **opts) as b:
        F0, T0, E0 = b.forces()
        d_pos, d_quat = _state4(hip, s)
        d_F = DeviceArray(ext_F.astype(ftype))
        d_T = DeviceArray(ext_T.astype(ftype))
        b.interaction_forces(d_pos.ptr, d_quat.ptr, d_F.ptr, d_T.ptr)
        b.sync()
        F1, T1, E1 = b.forces()
        U = b.energies()[0, 0]
    got_F = d_F.numpy()
    got_T = d_T.numpy()
    npt = np.float32 if precision == "mixed" else np.float64
    # the batch's forces of the same state, added in the working precision
    want_F = ext_F.astype(npt) + np.concatenate([F1, E1[:, None]], axis=1).astype(npt)
    assert np.array_equal(got_F, want_F)
    assert np.array_equal(got_T[:, :3], ext_T[:, :3].astype(npt) + T1.astype(npt))
    # the state given is the uploaded one up to the quaternion round trip
    assert vec_rel_err(F1, F0)[0] <= tol and vec_rel_err(T1, T0)[0] <= tol
    pair_F = got_F[:, :3].astype(np.float64) - ext_F[:, :3].astype(npt)
    pair_T = got_T[:, :3].astype(np.float64) - ext_T[:, :3].astype(npt)
    big_f, pp_f = vec_rel_err(pair_F, ref["F"])
    big_t, _ = vec_rel_err(pair_T, ref["T"])
    print(f"{name} {precision}: F {big_f:.2e}/{pp_f:.2e} T {big_t:.2e}")
    assert big_f <= 10 * tol and big_t <= 10 * tol   # FP32 additions onto O(1) values
    Esum = (got_F[:, 3].astype(np.float64) - ext_F[:, 3].astype(npt)).sum()
    scale = np.abs(ref["E8"]).sum()
    assert abs(Esum - U) <= 1e-5 * scale and abs(U - ref["E8"].sum()) <= 1e-5 * scale
    P = oracle_params(s, opts)
    assert abs(po.forces(P, s)["U"] - U) <= 1e-5 * scale


def test_moved_state_rebuilds_lists_on_callers_stream(hip):
    """Moving every particle by more than the skin between two calls
    rebuilds the lists: the forces of the moved state match the oracle.
    Work goes on a caller-created stream (default priority)."""
    from oracle import pyoracle as po
    s, opts, ref = load_golden("dense_md_excl_30C")
    P = oracle_params(s, opts)
    rng = np.random.default_rng(11)
    stream = Stream()
    with hip.HIPBatch(s, **opts) as b:
        reb0, _ = b.info()
        t = s.copy()
        for it in range(3):
            # a rigid shift of 0.3 (more than the skin) plus thermal-size noise
            t.pos = s.pos + 0.3 * (it + 1) + 0.02 * rng.normal(size=s.pos.shape)
            want = po.forces(P, t)
            d_pos, d_quat = _state4(hip, t)
            d_F = DeviceArray.zeros((s.N, 4), np.float32)
            d_T = DeviceArray.zeros((s.N, 4), np.float32)
            b.interaction_forces(d_pos.ptr, d_quat.ptr, d_F.ptr, d_T.ptr, stream.handle)
            stream.synchronize()
            F = d_F.numpy()[:, :3].astype(np.float64)
            T = d_T.numpy()[:, :3].astype(np.float64)
            big_f, pp_f = vec_rel_err(F, want["F"])
            big_t, pp_t = vec_rel_err(T, want["T"])
            print(f"move {it}: F {big_f:.2e}/{pp_f:.2e} T {big_t:.2e}/{pp_t:.2e}")
            assert big_f <= 1e-5 and big_t <= 1e-5 and pp_f <= 1e-5 and pp_t <= 1e-5
        reb1, _ = b.info()
    stream.close()
    assert reb1 - reb0 == 3


def test_rejects_pools_and_null_buffers(hip):
    from oxdna_amd import _lib
    s, opts, _ = load_golden("hairpin_md_50C")
    with hip.HIPBatch(s, **opts) as b:
        with pytest.raises(_lib.OxhError):
            b.interaction_forces(0, 0, 0, 0)
    with hip.HIPPool([64]) as pool:
        with pytest.raises(_lib.OxhError, match="slot pools"):
            pool.interaction_forces(1, 1, 1, 1)
