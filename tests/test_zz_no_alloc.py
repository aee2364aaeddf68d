"""The allocation-free force entry point (oxh_interaction_compute_forces_noalloc)
against the batch entry point it varies (oxh_interaction_compute_forces), through
the C ABI with device buffers from the HIP runtime (tests/_hipmem.py) as the
caller's memory.

Both paths run the same kernels with the same launch shape (8 lanes per
nucleotide, far pairs inline, in-place close kernel, all-pairs list build for
systems of at most 1024 nucleotides), so the additions agree bit for bit.
Two calls in a row share the library's scratch array, the second on a smaller
system than the first; both are checked, and against the FP64 oracle at the
force tests' bars (1e-5 of the largest force, FP32 form; 1e-9, FP64 form).
"""
import numpy as np
import pytest

from _common import load_golden, vec_rel_err
from _hipmem import DeviceArray, synchronize

pytestmark = pytest.mark.gpu

PINNED = dict(OXH_FORCE_LANES="8", OXH_FAR_LANES="0", OXH_PACK="0", OXH_EDGE="0", OXH_PERSIST="0")


def _state(backend, s):
    pos4 = np.zeros((s.N, 4))
    pos4[:, :3] = s.pos
    return DeviceArray(pos4), DeviceArray(backend.quat_from_frame(s.a1, s.a3))


@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_noalloc_matches_batch_entry_point(precision, monkeypatch):
    from oxdna_amd import _lib, backend
    assert _lib.device_count() >= 1, "no GPU visible"
    for k, v in PINNED.items():
        monkeypatch.setenv(k, v)
    ftype = np.float32 if precision == "mixed" else np.float64
    tol = 1e-5 if precision == "mixed" else 1e-9
    names = ["gas8_25C_0.5M", "bundle_md_30C"]   # 384 then 144 nucleotides
    cases = []
    rng = np.random.default_rng(17)
    for name in names:
        s, opts, ref = load_golden(name)
        preF = rng.normal(size=(s.N, 4))
        preT = rng.normal(size=(s.N, 4))
        d_pos, d_quat = _state(backend, s)
        # the existing entry point on a batch of the same system
        dF, dT = DeviceArray(preF.astype(ftype)), DeviceArray(preT.astype(ftype))
        with backend.HIPBatch(s, backend_precision=precision, **opts) as b:
            b.interaction_forces(d_pos.ptr, d_quat.ptr, dF.ptr, dT.ptr)
            b.sync()
        cases.append((name, s, opts, ref, preF, preT, d_pos, d_quat, dF.numpy(), dT.numpy()))
    # the allocation-free entry point: two calls in a row, the second smaller
    outs = []
    for name, s, opts, ref, preF, preT, d_pos, d_quat, _, _ in cases:
        dF, dT = DeviceArray(preF.astype(ftype)), DeviceArray(preT.astype(ftype))
        backend.interaction_forces_noalloc(s, d_pos.ptr, d_quat.ptr, dF.ptr, dT.ptr, backend_precision=precision, **opts)
        outs.append((dF, dT))
    synchronize()
    npt = np.float32 if precision == "mixed" else np.float64
    for (name, s, opts, ref, preF, preT, _, _, wantF, wantT), (dF, dT) in zip(cases, outs):
        gotF, gotT = dF.numpy(), dT.numpy()
        assert np.array_equal(gotF, wantF), f"{name}: forces differ from the batch entry point"
        assert np.array_equal(gotT, wantT), f"{name}: torques differ from the batch entry point"
        pair_F = gotF[:, :3].astype(np.float64) - preF[:, :3].astype(npt)
        pair_T = gotT[:, :3].astype(np.float64) - preT[:, :3].astype(npt)
        big_f, _ = vec_rel_err(pair_F, ref["F"])
        big_t, _ = vec_rel_err(pair_T, ref["T"])
        print(f"{name} {precision}: F {big_f:.2e} T {big_t:.2e}")
        assert big_f <= 10 * tol and big_t <= 10 * tol   # rounding of the additions onto O(1) values


def test_noalloc_rna_matches_batch_entry_point(monkeypatch):
    """The same for oxRNA2 systems (interaction in the descriptor, FP64
    terms): bit for bit against the batch entry point, and against the
    reference binary's fixtures of HB-free states at the FP64 bar."""
    from oxdna_amd import _lib, backend
    from test_gpu_rna import load_rna
    assert _lib.device_count() >= 1, "no GPU visible"
    for k, v in PINNED.items():
        monkeypatch.setenv(k, v)
    rng = np.random.default_rng(23)
    cases = []
    for name in ["rna_contacts_nohb", "rna_ff16_nohb"]:   # 432 then 16 nucleotides
        s, opts, ref = load_rna(name)
        preF, preT = rng.normal(size=(s.N, 4)), rng.normal(size=(s.N, 4))
        d_pos, d_quat = _state(backend, s)
        dF, dT = DeviceArray(preF), DeviceArray(preT)
        with backend.HIPBatch(s, backend_precision="double", **opts) as b:
            b.interaction_forces(d_pos.ptr, d_quat.ptr, dF.ptr, dT.ptr)
            b.sync()
        cases.append((name, s, opts, ref, preF, preT, d_pos, d_quat, dF.numpy(), dT.numpy()))
    outs = []
    for name, s, opts, ref, preF, preT, d_pos, d_quat, _, _ in cases:
        dF, dT = DeviceArray(preF), DeviceArray(preT)
        backend.interaction_forces_noalloc(s, d_pos.ptr, d_quat.ptr, dF.ptr, dT.ptr, backend_precision="double", **opts)
        outs.append((dF, dT))
    synchronize()
    for (name, s, opts, ref, preF, preT, _, _, wantF, wantT), (dF, dT) in zip(cases, outs):
        gotF, gotT = dF.numpy(), dT.numpy()
        assert np.array_equal(gotF, wantF) and np.array_equal(gotT, wantT), name
        big_f, _ = vec_rel_err(gotF[:, :3] - preF[:, :3], ref["F"])
        big_t, _ = vec_rel_err(gotT[:, :3] - preT[:, :3], ref["T"])
        print(f"{name}: F {big_f:.2e} T {big_t:.2e}")
        assert big_f <= 1e-8 and big_t <= 1e-8


def test_noalloc_rejects_oversize_system():
    from oxdna_amd import _lib, backend, systems
    assert _lib.device_count() >= 1, "no GPU visible"
    s = systems.duplex_gas(3, n_duplex=216, bp=24, density=1e-3, tc=25.0)   # 10,368 nucleotides
    assert s.N > 4096
    buf = DeviceArray.zeros((s.N, 4), np.float64)
    with pytest.raises(_lib.OxhError, match="at most 4096"):
        backend.interaction_forces_noalloc(s, buf.ptr, buf.ptr, buf.ptr, buf.ptr)
    assert not buf.numpy().any()
