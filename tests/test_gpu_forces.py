"""GPU parity of the oxDNA2 force kernel (through the C ABI) against the
reference's FP64 CPU backend (fixtures written by the reference binary)
and the FP64 oracle.

Tolerance (FP32 forces from FP64 positions): max |dF_i| over particles
<= 1e-5 of the largest force in the system, for forces and torques; the
per-particle error with a floor at 0.1 x RMS is reported and bounded by
1e-4.  Energies: each term within 1e-5 x sum|E_terms| + 1e-6 N.
"""
import numpy as np
import pytest

from _common import golden_names, load_golden, oracle_params, vec_rel_err

pytestmark = pytest.mark.gpu

F_TOL = 1e-5
F_TOL_PER_PARTICLE = 1e-4


@pytest.fixture(scope="module")
def hip():
    from oxdna_amd import _lib, backend
    assert _lib.device_count() >= 1, "no GPU visible"
    return backend


@pytest.mark.parametrize("name", golden_names())
def test_forces_vs_reference(hip, name):
    s, opts, ref = load_golden(name)
    with hip.HIPBatch(s, **opts) as b:
        F, T, Ep = b.forces()
        e8 = b.energy_split()
        U = b.energies()[0, 0]
    ef_big, ef_pp = vec_rel_err(F, ref["F"])
    et_big, et_pp = vec_rel_err(T, ref["T"])
    print(f"{name}: N={s.N} F {ef_big:.2e}/{ef_pp:.2e} T {et_big:.2e}/{et_pp:.2e}")
    assert ef_big <= F_TOL and et_big <= F_TOL
    assert ef_pp <= F_TOL_PER_PARTICLE and et_pp <= F_TOL_PER_PARTICLE
    scale = np.abs(ref["E8"]).sum() + 1e-6 * s.N
    assert np.all(np.abs(e8 - ref["E8"]) <= 1e-5 * scale), (e8, ref["E8"])
    assert abs(U - ref["E8"].sum()) <= 1e-5 * scale
    assert abs(Ep.sum() - U) <= 1e-5 * scale


@pytest.mark.parametrize("name", golden_names())
def test_stress_vs_reference(hip, name):
    """Pair virial against the reference's pressure observable (stress
    tensor x volume; zero velocities in the fixtures).  FP32 pair forces:
    each pair term carries ~1e-6 of the largest force, so the bound is
    1e-6 x N x max|F| (|r| ~ 1 for interacting pairs)."""
    s, opts, ref = load_golden(name)
    with hip.HIPBatch(s, **opts) as b:
        W, _ = b.stress()
    want = ref["stress7"][1:] * np.prod(s.box)
    scale = s.N * np.abs(ref["F"]).max()
    err = np.abs(W - want).max() / scale
    print(f"{name}: virial err {err:.2e} (|W| {np.abs(want).max():.3g}, scale {scale:.3g})")
    assert err <= 1e-6


@pytest.mark.parametrize("name", ["duplex20_md_20C", "bundle_md_30C", "gas8_25C_0.5M", "tangle2_45C_0.1M"])
def test_pair_set_exact(hip, name):
    """The GPU Verlet pair set equals the FP64 predicate's, bit for bit."""
    from oracle import pyoracle as po
    s, opts, _ = load_golden(name)
    P = oracle_params(s, opts)
    rv = P.rcut + 2 * 0.05
    with hip.HIPBatch(s, verlet_skin=0.05, **opts) as b:
        got = b.pairs()
    want = po.verlet_pairs(s.pos, s.n3, s.n5, s.box, rv * rv)
    assert got.shape == want.shape and np.array_equal(got, want)


def test_deterministic(hip):
    s, opts, _ = load_golden("bundle_md_30C")
    outs = []
    for _ in range(2):
        with hip.HIPBatch(s, thermostat="langevin", diff_coeff=2.5, seed=7, **opts) as b:
            b.run(200)
            st = b.state()
            F, T, _ = b.forces()
        outs.append((st["pos"], st["vel"], F, T))
    for a, c in zip(*outs):
        assert np.array_equal(a, c)


@pytest.mark.parametrize("lanes,far_lanes", [(1, 0), (8, 0), (4, 4), (8, 16), (32, 1)])
@pytest.mark.parametrize("name", ["bundle_md_30C", "tangle3_nohalf_dh", "gas8_25C_0.5M"])
def test_launch_shapes(hip, monkeypatch, name, lanes, far_lanes):
    """Every launch shape the autotuner may pick gives the same physics:
    far pairs inline or in their own kernel, any lanes per nucleotide."""
    monkeypatch.setenv("OXH_FORCE_LANES", str(lanes))
    monkeypatch.setenv("OXH_FAR_LANES", str(far_lanes))
    s, opts, ref = load_golden(name)
    with hip.HIPBatch(s, **opts) as b:
        b.compute_forces()
        F, T, _ = b.forces()
        e8 = b.energy_split()
        st = b.stats()
    assert st["lanes"] == lanes and st["far_lanes"] == far_lanes
    assert vec_rel_err(F, ref["F"])[0] <= F_TOL
    assert vec_rel_err(T, ref["T"])[0] <= F_TOL
    scale = np.abs(ref["E8"]).sum() + 1e-6 * s.N
    assert np.all(np.abs(e8 - ref["E8"]) <= 1e-5 * scale)
