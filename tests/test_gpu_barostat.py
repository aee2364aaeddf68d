"""The MC barostat (oxh_batch_barostat_move; the reference's VolumeMove /
MoleculeVolumeMove applied during MD, src/Backends/MCMoves/
VolumeMove.cpp:49-122, MoleculeVolumeMove.cpp) through the C ABI.

  - One move against the FP64 oracle: for given uniform numbers, the new
    box, the scaled positions (every particle, or every strand with its
    centre of mass), the energy change (the oracle's energy of the scaled
    state minus that of the old one) and the acceptance decision
    exp(-(dU + P dV - n T ln(V'/V)) / T) > u; a rejected move restores box
    and positions exactly.
  - The stationary law, closed form: for particles that never interact
    (single nucleotides farther apart than every cut-off in every box the
    chain visits) dU = 0, and isotropic moves uniform in the box side L
    sample p(L) ~ L^{3n} exp(-P L^3 / T), i.e. V ~ Gamma(n + 1/3, T / P):
    <V> = (n + 1/3) T / P, with n = N for the atomic move and the number of
    strands + 1 for the molecular one.
"""
import numpy as np
import pytest

from _common import load_golden, oracle_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from oxdna_amd import _lib, backend
    assert _lib.device_count() >= 1, "no GPU visible"
    return backend


def _expected(po, P, s, Pressure, delta_L, u4, isotropic, molecular):
    L0 = np.array(s.box, dtype=float)
    dL = delta_L * (np.array([u4[0]] * 3 if isotropic else u4[:3]) - 0.5)
    L1 = L0 + dL
    t = s.copy()
    if molecular:
        for k in np.unique(s.strand):
            m = s.strand == k
            t.pos[m] = s.pos[m] + s.pos[m].mean(axis=0) * (L1 / L0 - 1.0)
        n = len(np.unique(s.strand)) + 1
    else:
        t.pos = s.pos * (L1 / L0)
        n = s.N
    t.box = L1
    U0 = po.forces(P, s)["U"]
    U1 = po.forces(P, t)["U"]
    V0, V1 = np.prod(L0), np.prod(L1)
    x = -((U1 - U0) + Pressure * (V1 - V0) - n * s.T * np.log(V1 / V0)) / s.T
    return t, U1 - U0, x


@pytest.mark.parametrize("precision,tol", [("mixed", 1e-5), ("double", 1e-9)])
@pytest.mark.parametrize("isotropic,molecular", [(True, False), (False, False), (True, True), (False, True)])
def test_one_move_against_the_oracle(hip, precision, tol, isotropic, molecular):
    from oracle import pyoracle as po
    s, opts, _ = load_golden("gas8_25C_0.5M")
    P = oracle_params(s, opts)
    rng = np.random.default_rng(3)
    seen = []
    # strands keep their shape in a molecular move; an atomic move
    # stretches every bond, so its steps are smaller
    dl = 1.5 if molecular else 0.3
    with hip.HIPBatch(s, backend_precision=precision, **opts) as b:
        for attempt in range(40):
            if len(seen) >= 8 and len(set(seen)) == 2:
                break
            u4 = rng.random(4)
            t, dU, x = _expected(po, P, s, 2e-4, dl, u4, isotropic, molecular)
            if abs(x - np.log(u4[3])) < 1e-2:
                continue   # too close to the decision for FP32 energies
            st0 = b.state()
            acc, dE = b.barostat_move(2e-4, dl, u4, isotropic, molecular)
            assert abs(dE - dU) <= 1e2 * tol * max(1.0, abs(po.forces(P, s)["U"])), (dE, dU)
            assert acc == (np.exp(x) > u4[3])
            st = b.state()
            if acc:
                assert np.allclose(b.box(), t.box, rtol=0, atol=1e-12)
                assert np.abs(st["pos"] - t.pos).max() <= 1e-12 * np.abs(t.pos).max()
                s = t
            else:
                assert np.array_equal(b.box(), np.array(s.box, dtype=float))
                assert np.array_equal(st["pos"], st0["pos"])
            seen.append(acc)
        F, T, _ = b.forces()
        ref = po.forces(P, s)
        assert np.abs(F - ref["F"]).max() <= 10 * tol * np.abs(ref["F"]).max()
    assert set(seen) == {True, False}


@pytest.mark.parametrize("molecular", [False, True])
def test_ideal_gas_volume_follows_the_closed_form(hip, molecular):
    from oxdna_amd import systems
    # 8 single nucleotides on a 2 x 2 x 2 lattice of spacing L / 2: the
    # chain's boxes stay above 16 units, so every pair stays beyond the
    # cut-offs and dU = 0
    N, T, L = 8, systems.celsius(25.0), 40.0
    n = N if not molecular else N + 1
    Pressure = (n + 1.0 / 3.0) * T / L ** 3
    b0 = systems._Builder()
    for x in (0.25, 0.75):
        for y in (0.25, 0.75):
            for z in (0.25, 0.75):
                R = np.eye(3)
                b0.add_strand([np.array([x, y, z]) * L], [R[:, 0]], [R[:, 2]], [0])
    s = b0.build([L, L, L], "gas", T=T, salt=0.5)
    rng = np.random.default_rng(11)
    vols = []
    with hip.HIPBatch(s, backend_precision="double") as b:
        for it in range(3000):
            acc, dE = b.barostat_move(Pressure, 12.0, rng.random(4), True, molecular)
            assert dE == 0.0
            vols.append(np.prod(b.box()))
    v = np.array(vols[200:])
    assert v.min() > 16.0 ** 3
    # standard error from 20 blocks
    blocks = v[: len(v) // 20 * 20].reshape(20, -1).mean(axis=1)
    se = blocks.std(ddof=1) / np.sqrt(len(blocks))
    want = (n + 1.0 / 3.0) * T / Pressure
    z = (v.mean() - want) / se
    print(f"molecular={molecular}: <V> {v.mean():.1f} want {want:.1f} (se {se:.1f}, z {z:+.2f})")
    assert abs(z) <= 4.0
    # the other count is excluded: n -/+ 1 gives a mean 12 % away
    other = (n + (1 if not molecular else -1) + 1.0 / 3.0) * T / Pressure
    assert abs(v.mean() - other) > 4.0 * se
