"""Pins the FP64 CPU oracle (oracle/oxdna2_oracle.cpp).

The oracle is the checker of every GPU parity test, so it is checked here
against things that do not share its formulas:
  - the reference's own fixture test/DNA/FORCE_FIELD/AVG_SEQ/reference.dat
    (copied to tests/golden/avg_seq_reference.dat; meshed f4, so agreement
    is to the mesh error, 5e-6 per nucleotide);
  - energies, forces and torques written by the reference binary built from
    source (tests/golden/*.npz, make_golden.py), to 1e-12;
  - F = -grad U and T = -dU/dphi by central finite differences of the
    oracle's own total energy (a sign or index slip in any force or torque
    term fails these);
  - Newton's third law and conservation of angular momentum;
  - the Verlet predicate against brute force, the Hilbert key against the
    defining properties of a Hilbert curve, Philox against published
    known-answer vectors, the integrator against NVE energy conservation
    and time reversibility.
"""
import os

import numpy as np
import pytest

from _common import GOLDEN, golden_names, load_golden, oracle_params
from oracle import pyoracle as po
from oxdna_amd import io, systems


@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_reference_binary(name):
    s, opts, ref = load_golden(name)
    r = po.forces(oracle_params(s, opts), s)
    assert np.all(np.abs(r["E8"] - ref["E8"]) <= 1e-12 * s.N)
    fs = np.abs(ref["F"]).max()
    ts = np.abs(ref["T"]).max()
    assert np.abs(r["F"] - ref["F"]).max() <= 1e-11 * fs
    assert np.abs(r["T"] - ref["T"]).max() <= 1e-11 * ts
    assert abs(r["Epart"].sum() - r["U"]) <= 1e-12 * s.N


@pytest.mark.parametrize("name", golden_names())
def test_oracle_virial_matches_reference_pressure(name):
    """Pair virial against the reference's pressure observable with its
    stress tensor (src/Observables/Pressure.cpp, printed with 8 significant
    digits; velocities are zero in the fixtures, so the observable holds the
    virial / V only)."""
    s, opts, ref = load_golden(name)
    W = po.virial(oracle_params(s, opts), s) / np.prod(s.box)
    st = ref["stress7"]
    scale = np.abs(st[1:]).max()
    assert np.all(np.abs(W - st[1:]) <= 2e-8 * scale + 1e-14), (W, st[1:])
    assert abs(W[:3].sum() / 3 - st[0]) <= 2e-8 * scale + 1e-14


def test_oracle_matches_reference_fixture_file():
    """reference.dat of the reference's FORCE_FIELD/AVG_SEQ test: DNA2
    (meshed f4), 16 nt, 20 C, 1.0 M, columns per nucleotide."""
    btype, n3, n5, strand = io.read_topology(os.path.join(GOLDEN, "ff_init.top"))
    step, box, pos, a1, a2, a3, vel, L = io.read_conf(os.path.join(GOLDEN, "ff_init.dat"), len(btype))
    s = systems.System(pos=pos, a1=a1, a3=a3, btype=btype, n3=n3, n5=n5, strand=strand, box=box)
    want = np.array([float(x) for x in open(os.path.join(GOLDEN, "avg_seq_reference.dat")).read().split()])
    r = po.forces(po.Params(T=io.temperature("20C"), salt=1.0), s)
    got = r["E8"] / s.N
    # the fixture is printed with 6 decimals; the mesh moves HB by ~1.3e-6
    assert np.all(np.abs(got - want) <= 5e-6), (got, want)


def _energy(P, s):
    return po.forces(P, s)["U"]


def _rotate_particle(s, i, axis, ang):
    R = systems.rotation_matrix(axis, ang)
    out = s.copy()
    out.a1[i] = R @ s.a1[i]
    out.a3[i] = R @ s.a3[i]
    return out


@pytest.mark.parametrize("name", golden_names())
def test_oracle_hb_count_matches_reference(name):
    """Base pairs (H-bond energy below HB_CUTOFF = -0.1, src/defs.h:62)
    against the reference's hb_list observable with only_count = true
    (src/Observables/HBList.cpp:141-164)."""
    s, opts, ref = load_golden(name)
    assert po.count_hb(oracle_params(s, opts), s) == ref["hb_count"]


def test_fixtures_cover_every_term():
    """Every term of the energy split is nonzero in at least three
    fixtures, bonded excluded volume and coaxial stacking included."""
    nz = np.array([np.abs(load_golden(n)[2]["E8"]) > 1e-9 for n in golden_names()])
    assert np.all(nz.sum(axis=0) >= 3), nz.sum(axis=0)


@pytest.mark.parametrize("name", ["hairpin_md_50C", "nicked_md_30C", "tangle3_nohalf_dh", "bundle_md_30C",
                                  "compressed_20C", "tangle4_dense_30C", "hot_md_bexcl_95C", "hot_md_cxst_95C",
                                  "mbf_md_20C", "hbmult_custom_25C", "seqdep_hairpin_45C"])
def test_forces_and_torques_are_energy_gradients(name):
    s, opts, _ = load_golden(name)
    P = oracle_params(s, opts)
    r = po.forces(P, s)
    rng = np.random.default_rng(0)
    h = 1e-5
    # central differences err by h^2/6 times the third derivative, which is
    # large in the Lennard-Jones cores of the packed fixtures
    tol = 1e-5 if np.abs(r["F"]).max() < 100 else 1e-4
    # a random sample plus the four most strongly pushed particles (those
    # in excluded-volume contact in the packed fixtures)
    picks = set(rng.choice(s.N, size=min(6, s.N), replace=False).tolist())
    picks |= set(np.argsort(-np.linalg.norm(r["F"], axis=1))[:4].tolist())
    for i in sorted(picks):
        for d in range(3):
            sp = s.copy()
            sm = s.copy()
            sp.pos[i, d] += h
            sm.pos[i, d] -= h
            fd = -(_energy(P, sp) - _energy(P, sm)) / (2 * h)
            assert abs(fd - r["F"][i, d]) <= tol * max(1.0, abs(r["F"][i, d])), (i, d, fd, r["F"][i, d])
        axes = (s.a1[i], np.cross(s.a3[i], s.a1[i]), s.a3[i])
        for k in range(3):
            up = _energy(P, _rotate_particle(s, i, axes[k], h))
            um = _energy(P, _rotate_particle(s, i, axes[k], -h))
            fd = -(up - um) / (2 * h)
            assert abs(fd - r["T"][i, k]) <= tol * max(1.0, abs(r["T"][i, k])), (i, k, fd, r["T"][i, k])


@pytest.mark.parametrize("name", ["duplex20_md_20C", "hairpin_md_50C", "nicked_md_30C"])
def test_newton_third_law_and_angular_momentum(name):
    s, opts, _ = load_golden(name)
    r = po.forces(oracle_params(s, opts), s)
    F = r["F"]
    a2 = np.cross(s.a3, s.a1)
    T_lab = r["T"][:, 0:1] * s.a1 + r["T"][:, 1:2] * a2 + r["T"][:, 2:3] * s.a3
    assert np.abs(F.sum(axis=0)).max() <= 1e-10 * np.abs(F).max()
    # compact systems: nearest images equal plain differences
    tot = np.cross(s.pos, F).sum(axis=0) + T_lab.sum(axis=0)
    assert np.abs(tot).max() <= 1e-9 * (np.abs(F).max() * np.abs(s.pos).max() + np.abs(T_lab).max())


def test_verlet_predicate_cells_match_brute_force():
    s = systems.duplex_gas(3, n_duplex=27, bp=40, density=0.02)   # N = 2160 > cell threshold
    rv = 3.357
    got = po.verlet_pairs(s.pos, s.n3, s.n5, s.box, rv * rv)
    d = s.pos[None, :, :] - s.pos[:, None, :]
    d -= s.box * np.rint(d / s.box)
    r2 = (d ** 2).sum(-1)
    i, j = np.nonzero(np.triu(r2 < rv * rv, 1))
    keep = ~((s.n3[i] == j) | (s.n5[i] == j))
    want = np.stack([i[keep], j[keep]], 1)
    # pairs within rounding of the threshold may differ between numpy and the
    # fused-multiply-add predicate; there are none at this seed
    assert np.all(np.abs(r2[i, j] - rv * rv) > 1e-9)
    assert np.array_equal(got, want[np.lexsort((want[:, 1], want[:, 0]))])


@pytest.mark.parametrize("bits", [1, 2, 3, 4])
def test_hilbert_key_is_a_hilbert_curve(bits):
    n = 1 << bits
    keys = {}
    for x in range(n):
        for y in range(n):
            for z in range(n):
                keys[po.hilbert_key(x, y, z, bits)] = (x, y, z)
    assert sorted(keys) == list(range(n ** 3))          # bijection onto [0, n^3)
    path = np.array([keys[k] for k in range(n ** 3)])
    steps = np.abs(np.diff(path, axis=0)).sum(1)
    assert np.all(steps == 1)                            # consecutive cells share a face


def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32-10
    cases = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in cases:
        assert list(po.philox(ctr, key)) == want


def test_oracle_integrator_conserves_energy_and_is_reversible():
    s = systems.duplex_20bp(4)
    P = po.Params(T=s.T, salt=0.5)
    r = po.forces(P, s)
    F, T = r["F"].copy(), r["T"].copy()
    rng = np.random.default_rng(1)
    vel = rng.normal(scale=np.sqrt(s.T), size=(s.N, 3))
    L = rng.normal(scale=np.sqrt(s.T), size=(s.N, 3))
    start = (s.pos.copy(), s.a1.copy(), s.a3.copy())
    E0 = r["U"] + 0.5 * (vel ** 2).sum() + 0.5 * (L ** 2).sum()
    for _ in range(300):
        U = po.md_step(P, s, 0.003, F, T, vel, L)
    E1 = U + 0.5 * (vel ** 2).sum() + 0.5 * (L ** 2).sum()
    assert abs(E1 - E0) / s.N < 2e-3
    vel *= -1
    L *= -1
    for _ in range(300):
        po.md_step(P, s, 0.003, F, T, vel, L)
    assert np.abs(s.pos - start[0]).max() < 1e-7
    assert np.abs(s.a1 - start[1]).max() < 1e-7
