"""GPU tests of the oxRNA / oxRNA2 pair forces (k_forces_rna) through the C
ABI, against the reference binary's fixtures (tests/golden/rna/, written by
tests/golden/make_rna_golden.py).

  - states without hydrogen bonds (the "_nohb" twins, the collapsed single
    strands): forces, torques, the energy split and the stress tensor equal
    the reference binary's to FP64 rounding (1e-9 relative), in the
    all-FP64 form; the mixed form (angular factors in FP32, FP32 outputs)
    at the oxDNA2 mixed form's bar (1e-5 of the largest force);
  - states with hydrogen bonds: every energy term but HB equals the
    reference's to FP64 rounding; the HB energy and the forces differ by
    the reference CPU backend's cubic-mesh interpolation of the HB f4
    factors (RNAInteraction.cpp:312-324), which the GPU evaluates
    analytically as the reference's GPU kernel does (CUDA_RNA.cuh:743-761):
    bars of 5e-5 of the HB energy and 2e-3 of the largest force and torque,
    about ten times the largest departure measured on MI355X (5.4e-6 and
    3.3e-4).  The HB forces and torques themselves
    are pinned to the HB energy by central differences of the GPU's own
    energy (translations and rotations of single nucleotides);
  - the reference's published split of its oxRNA force-field test
    (test/RNA/FORCE_FIELD/AVG_SEQ/reference.dat, 6 decimals);
  - every launch shape gives the same forces; an NVE run conserves energy.
"""
import os

import numpy as np
import pytest

from _common import GOLDEN, vec_rel_err

pytestmark = pytest.mark.gpu

RNA_DIR = os.path.join(GOLDEN, "rna")
NAMES = sorted(f[:-4] for f in os.listdir(RNA_DIR) if f.endswith(".npz"))


def load_rna(name):
    from oxdna_amd import systems
    z = np.load(os.path.join(RNA_DIR, name + ".npz"))
    s = systems.System(pos=z["pos"].copy(), a1=z["a1"].copy(), a3=z["a3"].copy(), btype=z["btype"].copy(),
                       n3=z["n3"].copy(), n5=z["n5"].copy(), strand=z["strand"].copy(), box=z["box"].copy(),
                       T=float(z["T"]), salt=float(z["salt"]), name=name)
    opts = dict(interaction_type=str(z["interaction"]), dh_half_charged_ends=bool(z["half_charged_ends"]),
                dh_lambda=float(z["dh_lambda"]), dh_strength=float(z["dh_strength"]),
                debye_huckel_rhigh=float(z["dh_rhigh"]), max_backbone_force=float(z["max_backbone_force"]),
                max_backbone_force_far=float(z["max_backbone_force_far"]))
    ref = dict(E8=z["E8"].copy(), F=z["F"].copy(), T=z["Tbody"].copy(), stress7=z["stress7"].copy(),
               hb_count=int(z["hb_count"]))
    # oxRNA without Debye-Hueckel splits its energy into seven terms
    if len(ref["E8"]) == 7:
        ref["E8"] = np.append(ref["E8"], 0.0)
    if "E8_published" in z:
        ref["published"] = z["E8_published"].copy()
    return s, opts, ref


@pytest.fixture(scope="module")
def hip():
    from oxdna_amd import _lib, backend
    assert _lib.device_count() >= 1, "no GPU visible"
    return backend


def _eval(hip, s, opts, precision="double"):
    with hip.HIPBatch(s, backend_precision=precision, **opts) as b:
        F, T, Ep = b.forces()
        e8 = b.energy_split()
        W, _ = b.stress()
        hb = int(b.hb_count()[0])
        lanes = b.stats()["lanes"]
    return F, T, Ep, e8, W, hb, lanes


@pytest.mark.parametrize("name", NAMES)
def test_rna_fp64_vs_reference(hip, name):
    s, opts, ref = load_rna(name)
    F, T, Ep, e8, W, hb, lanes = _eval(hip, s, opts)
    ef = vec_rel_err(F, ref["F"])[0]
    et = vec_rel_err(T, ref["T"])[0]
    scale = np.abs(ref["E8"]).sum() + 1e-6 * s.N
    ehb = abs(e8[4] - ref["E8"][4])
    print(f"{name}: lanes {lanes} F {ef:.2e} T {et:.2e} HB {ehb:.2e} of {abs(ref['E8'][4]):.3e} hb {hb}/{ref['hb_count']}")
    other = [0, 1, 2, 3, 5, 6, 7]
    assert np.all(np.abs(e8[other] - ref["E8"][other]) <= 1e-10 * scale)
    assert abs(Ep.sum() - e8.sum()) <= 1e-10 * scale
    want = ref["stress7"][1:] * np.prod(s.box)
    if ref["E8"][4] == 0.0:
        assert e8[4] == 0.0 and hb == 0
        assert ef <= 1e-9 and et <= 1e-9
        # the reference prints the pressure observable with 8 significant digits
        assert np.all(np.abs(W - want) <= 2e-8 * np.abs(want).max() + 1e-12)
    else:
        assert ehb <= 5e-5 * abs(ref["E8"][4])
        assert ef <= 2e-3 and et <= 2e-3
        assert abs(hb - ref["hb_count"]) <= max(1, ref["hb_count"] // 50)
        print(f"  stress {np.abs(W - want).max() / np.abs(want).max():.2e}")
        assert np.all(np.abs(W - want) <= 5e-3 * np.abs(want).max() + 1e-12)


@pytest.mark.parametrize("name", ["rna_contacts_nohb", "rna_nicked24_nohb", "rna_tangle12_nohb"])
def test_rna_mixed_vs_reference(hip, name):
    """The mixed form: the angular factors (acos, sin, f4) in FP32, every
    other quantity in FP64, the results rounded into FP32 outputs."""
    s, opts, ref = load_rna(name)
    F, T, Ep, e8, W, hb, _ = _eval(hip, s, opts, precision="mixed")
    ef = vec_rel_err(F, ref["F"])[0]
    et = vec_rel_err(T, ref["T"])[0]
    scale = np.abs(ref["E8"]).sum() + 1e-6 * s.N
    de = np.abs(e8 - ref["E8"]).max() / scale
    print(f"{name}: mixed F {ef:.2e} T {et:.2e} split {de:.2e}")
    assert ef <= 1e-5 and et <= 1e-5
    assert de <= 1e-6


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16])
def test_rna_launch_shapes(hip, monkeypatch, lanes):
    monkeypatch.setenv("OXH_FORCE_LANES", str(lanes))
    s, opts, ref = load_rna("rna_contacts_nohb")
    F, T, Ep, e8, W, hb, got = _eval(hip, s, opts)
    assert got == lanes
    assert vec_rel_err(F, ref["F"])[0] <= 1e-9 and vec_rel_err(T, ref["T"])[0] <= 1e-9


def test_rna_published_split(hip):
    """The reference's own oxRNA2 test: its published per-nucleotide split
    (6 decimals).  The HB column carries the CPU mesh's departure too."""
    s, opts, ref = load_rna("rna_ff16")
    _, _, _, e8, _, _, _ = _eval(hip, s, opts)
    got = e8 / s.N
    print("published", ref["published"], "got", np.round(got, 6))
    other = [0, 1, 2, 3, 5, 6, 7]
    assert np.all(np.abs(got[other] - ref["published"][other]) <= 6e-7)
    assert abs(got[4] - ref["published"][4]) <= 5e-5 * abs(ref["published"][4]) + 6e-7


def _quat_rotate(a1, a3, axis, ang):
    from oxdna_amd import systems
    R = systems.rotation_matrix(axis, ang)
    return R @ a1, R @ a3


@pytest.mark.parametrize("name", ["rna_ff16", "rna_contacts_45C"])
def test_rna_forces_are_energy_gradients(hip, name):
    """Central differences of the GPU's FP64 energy (HB included, analytic
    f4) against its forces and torques, on the nucleotides with the
    largest HB energy share."""
    s, opts, ref = load_rna(name)
    with hip.HIPBatch(s, backend_precision="double", **opts) as b:
        F, T, Ep = b.forces()
        order = np.argsort(Ep)[:3]   # the most negative energy shares
        h = 1e-5

        def energy(pos, a1, a3):
            b.set_state(0, pos=pos, a1=a1, a3=a3)
            b.compute_forces()
            return float(b.energy_split().sum())

        for i in order:
            for k in range(3):
                p1, p2 = s.pos.copy(), s.pos.copy()
                p1[i, k] += h
                p2[i, k] -= h
                fd = -(energy(p1, s.a1, s.a3) - energy(p2, s.a1, s.a3)) / (2 * h)
                assert abs(fd - F[i, k]) <= 1e-5 * (1 + abs(F[i, k])), (i, k, fd, F[i, k])
            a2 = np.cross(s.a3[i], s.a1[i])
            Tlab = T[i, 0] * s.a1[i] + T[i, 1] * a2 + T[i, 2] * s.a3[i]
            for k in range(3):
                axis = np.eye(3)[k]
                A1, A3 = s.a1.copy(), s.a3.copy()
                B1, B3 = s.a1.copy(), s.a3.copy()
                A1[i], A3[i] = _quat_rotate(s.a1[i], s.a3[i], axis, h)
                B1[i], B3[i] = _quat_rotate(s.a1[i], s.a3[i], axis, -h)
                fd = -(energy(s.pos, A1, A3) - energy(s.pos, B1, B3)) / (2 * h)
                assert abs(fd - Tlab[k]) <= 1e-5 * (1 + abs(Tlab[k])), (i, k, fd, Tlab[k])


def test_rna_nve_conserves_energy(hip):
    """1000 NVE steps of a 40-bp oxRNA2 duplex in the all-FP64 form: the
    total energy drifts by less than 1e-4 per nucleotide."""
    from oxdna_amd import systems
    rng = np.random.default_rng(2)
    b0 = systems._Builder()
    systems.add_rna_duplex(b0, rng.integers(0, 4, 40), (15.0, 15.0, 15.0))
    s = b0.build([30.0, 30.0, 30.0], "rna40", T=systems.celsius(25.0), salt=1.0)
    s.vel = rng.normal(scale=np.sqrt(s.T), size=(s.N, 3))
    s.L = rng.normal(scale=np.sqrt(s.T), size=(s.N, 3))
    with hip.HIPBatch(s, backend_precision="double", interaction_type="RNA2", dt=0.002) as b:
        e0 = b.energies()[0]
        b.run(1000)
        e1 = b.energies()[0]
        err = b.info()[1]
    U0, K0 = e0[0], e0[2] + e0[3]
    U1, K1 = e1[0], e1[2] + e1[3]
    drift = abs((U1 + K1) - (U0 + K0)) / s.N
    print(f"E/N {(U0 + K0) / s.N:.6f} -> {(U1 + K1) / s.N:.6f}, drift {drift:.2e}")
    assert err == 0
    assert drift < 1e-4


def test_rna_refuses_cell_lists(hip):
    """oxRNA systems large enough for cell lists are refused at creation
    (the oxRNA path runs on systems of at most 1,024 nucleotides)."""
    from oxdna_amd import systems
    rng = np.random.default_rng(4)
    b0 = systems._Builder()
    for k in range(40):
        systems.add_rna_duplex(b0, rng.integers(0, 4, 16), ((k % 8) * 3.0 + 1.5, (k // 8) * 3.0 + 1.5, 10.0))
    s = b0.build([24.0, 24.0, 20.0], "rna1280", T=systems.celsius(37.0), salt=1.0)
    assert s.N == 1280
    with pytest.raises(Exception, match="without cell lists"):
        hip.HIPBatch(s, interaction_type="RNA2")


def test_rna_batch_of_systems(hip):
    """A batch of independent oxRNA2 systems (one launch for all) gives
    each system the forces it has alone (to FP64 rounding: the launch
    shapes, and so the order of the sums, may differ)."""
    from oxdna_amd import systems
    rs = systems.rna_duplex_set(3, n_sys=3, per_side=2, bp=12, box=12.0)
    with hip.HIPBatch(rs, backend_precision="double", interaction_type="RNA2") as b:
        together = [b.forces(k) for k in range(3)]
    for k, s in enumerate(rs):
        with hip.HIPBatch(s, backend_precision="double", interaction_type="RNA2") as c:
            F, T, E = c.forces()
        assert vec_rel_err(together[k][0], F)[0] <= 1e-12 and vec_rel_err(together[k][1], T)[0] <= 1e-12



@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_driver_runs_reference_rna_test_input(hip, tmp_path, precision):
    """The reference's own oxRNA2 test (test/RNA/FORCE_FIELD: its input file,
    topology in the new format with type=RNA, configuration; kept under
    tests/golden/rna/ff16_input) run unchanged by oxdna_amd.driver: the
    split_energy.dat it writes matches the reference's published
    reference.dat (6 decimals; the HB column within the CPU mesh's
    departure)."""
    import shutil
    from oxdna_amd import driver
    src = os.path.join(RNA_DIR, "ff16_input")
    (tmp_path / "run").mkdir()
    for f in ("init.top", "init.dat"):
        shutil.copy(os.path.join(src, f), tmp_path / f)
    shutil.copy(os.path.join(src, "run", "input"), tmp_path / "run" / "input")
    driver.run(str(tmp_path / "run" / "input"), [f"backend_precision = {precision}"])
    got = np.loadtxt(tmp_path / "run" / "split_energy.dat", ndmin=2)[0]
    want = np.loadtxt(os.path.join(src, "run", "reference.dat"))
    print(precision, "got", got, "want", want)
    assert got.shape == want.shape
    other = [0, 1, 2, 3, 5, 6, 7]
    assert np.all(np.abs(got[other] - want[other]) <= 1.01e-6)
    assert abs(got[4] - want[4]) <= 5e-5 * abs(want[4]) + 1.01e-6


def test_rna_pair_quantities(hip):
    """FFS order-parameter quantities of oxRNA2 pairs: the HB energies of
    every non-bonded pair sum to the split's HB term, base-base distances
    follow the sites (0.4f a1) under the nearest image, and the oxDNA-only
    near-HB criterion is refused."""
    s, opts, ref = load_rna("rna_contacts_45C")
    L = s.box
    d = s.pos[:, None, :] - s.pos[None, :, :]
    d -= L * np.rint(d / L)
    close = np.linalg.norm(d, axis=2) < 2.0
    bonded = np.zeros_like(close)
    for i in range(s.N):
        for j in (s.n3[i], s.n5[i]):
            if j >= 0:
                bonded[i, j] = True
    ii, jj = np.nonzero(np.triu(close & ~bonded, 1))
    pairs = np.stack([ii, jj], axis=1)
    with hip.HIPBatch(s, backend_precision="double", **opts) as b:
        e8 = b.energy_split()
        hb = b.pair_quantities("hb_energy", pairs)
        other = pairs[s.strand[ii] != s.strand[jj]]
        dist = b.pair_quantities("base_distance", other)
        with pytest.raises(Exception):
            b.pair_quantities("near_hb", pairs[:4])
    print(f"sum of pair HB {hb.sum():.12f}, split HB {e8[4]:.12f}, {len(pairs)} pairs")
    assert abs(hb.sum() - e8[4]) <= 1e-10 * abs(e8[4])
    pb = float(np.float32(0.4))   # RNA_POS_BASE = 0.4f (src/Interactions/rna_model.h:315)
    bi, bj = s.pos[other[:, 0]] + pb * s.a1[other[:, 0]], s.pos[other[:, 1]] + pb * s.a1[other[:, 1]]
    db = bj - bi
    db -= L * np.rint((s.pos[other[:, 1]] - s.pos[other[:, 0]]) / L)
    dd = np.abs(dist - np.linalg.norm(db, axis=1)).max()
    print(f"base distances: {len(other)} pairs, largest difference {dd:.2e}")
    assert dd <= 1e-12
