"""GPU tests of the all-FP64 force mode (backend_precision = double) and of
the integrator against the FP64 oracle, through the C ABI.

  - forces, torques, energies and the stress tensor of every fixture equal
    the reference binary's to FP64 rounding (1e-9 relative: the GPU fuses
    multiply-adds and uses its own libm, so bits differ);
  - NVE trajectories of the GPU (FP64 terms) follow the oracle's integrator
    (oracle/oxdna2_oracle.cpp, pinned by energy conservation and time
    reversibility) step for step; with FP32 terms they stay within what the
    FP32 force error allows.
"""
import numpy as np
import pytest

from _common import golden_names, load_golden, vec_rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from oxdna_amd import _lib, backend
    assert _lib.device_count() >= 1, "no GPU visible"
    return backend


@pytest.mark.parametrize("name", golden_names())
def test_fp64_forces_vs_reference(hip, name):
    s, opts, ref = load_golden(name)
    with hip.HIPBatch(s, backend_precision="double", **opts) as b:
        F, T, Ep = b.forces()
        e8 = b.energy_split()
        W, _ = b.stress()
        hb = int(b.hb_count()[0])
    assert hb == ref["hb_count"], (hb, ref["hb_count"])
    ef = vec_rel_err(F, ref["F"])[0]
    et = vec_rel_err(T, ref["T"])[0]
    print(f"{name}: FP64 F {ef:.2e} T {et:.2e}")
    assert ef <= 1e-9 and et <= 1e-9
    scale = np.abs(ref["E8"]).sum() + 1e-6 * s.N
    assert np.all(np.abs(e8 - ref["E8"]) <= 1e-10 * scale)
    want = ref["stress7"][1:] * np.prod(s.box)
    # the reference prints the pressure observable with 8 significant digits
    assert np.all(np.abs(W - want) <= 2e-8 * np.abs(want).max() + 1e-12)


@pytest.mark.parametrize("lanes,far_lanes", [(8, 0), (4, 8), (2, 16)])
@pytest.mark.parametrize("name", ["bundle_md_30C", "gas8_25C_0.5M", "tangle3_nohalf_dh"])
def test_fp64_launch_shapes(hip, monkeypatch, name, lanes, far_lanes):
    """All-FP64 far pairs inline or in their own kernel (k_forces_far_dp),
    any lanes: the same FP64 bar, energy split and virial included."""
    monkeypatch.setenv("OXH_FORCE_LANES", str(lanes))
    monkeypatch.setenv("OXH_FAR_LANES", str(far_lanes))
    s, opts, ref = load_golden(name)
    with hip.HIPBatch(s, backend_precision="double", **opts) as b:
        assert b.stats()["far_lanes"] == far_lanes
        b.compute_forces()
        F, T, Ep = b.forces()
        e8 = b.energy_split()
        W, _ = b.stress()
    assert vec_rel_err(F, ref["F"])[0] <= 1e-9 and vec_rel_err(T, ref["T"])[0] <= 1e-9
    scale = np.abs(ref["E8"]).sum() + 1e-6 * s.N
    assert np.all(np.abs(e8 - ref["E8"]) <= 1e-10 * scale)
    assert abs(Ep.sum() - ref["E8"].sum()) <= 1e-10 * scale
    want = ref["stress7"][1:] * np.prod(s.box)
    assert np.all(np.abs(W - want) <= 2e-8 * np.abs(want).max() + 1e-12)


@pytest.mark.parametrize("precision,tol", [("double", 1e-9), ("mixed", 1e-5)])
def test_nve_trajectory_follows_oracle(hip, precision, tol):
    from oracle import pyoracle as po
    from oxdna_amd import systems
    s = systems.duplex_20bp(4)
    rng = np.random.default_rng(1)
    s.vel = rng.normal(scale=np.sqrt(s.T), size=(s.N, 3))
    s.L = rng.normal(scale=np.sqrt(s.T), size=(s.N, 3))
    P = po.Params(T=s.T, salt=0.5)
    o = s.copy()
    r = po.forces(P, o)
    F, T = r["F"].copy(), r["T"].copy()
    vel, L = s.vel.copy(), s.L.copy()
    steps = 50
    for _ in range(steps):
        po.md_step(P, o, 0.003, F, T, vel, L)
    with hip.HIPBatch(s, backend_precision=precision, thermostat="no", dt=0.003, salt_concentration=0.5) as b:
        b.run(steps)
        st = b.state()
    dpos = np.abs(st["pos"] - o.pos).max()
    dvel = np.abs(st["vel"] - vel).max()
    da1 = np.abs(st["a1"] - o.a1).max()
    print(f"{precision}: |dpos| {dpos:.2e} |dvel| {dvel:.2e} |da1| {da1:.2e}")
    assert dpos <= tol and da1 <= tol and dvel <= 20 * tol
