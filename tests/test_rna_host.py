"""CPU tests of the oxRNA host side: interaction selection, input keys,
the oxRNA duplex generator and the consistency of the oxRNA fixtures."""
import os

import numpy as np
import pytest

from _common import GOLDEN

RNA_DIR = os.path.join(GOLDEN, "rna")


def test_interaction_field_of_the_descriptor():
    from oxdna_amd import _lib, backend, systems
    s = systems.duplex_20bp(0)
    keep = []
    assert backend.make_desc(s, keep).interaction == 0
    assert backend.make_desc(s, keep, interaction_type="RNA2").interaction == 1
    assert backend.make_desc(s, keep, interaction_type="RNA").interaction == 2
    assert backend.make_desc(s, keep, interaction_type="DNA2_nomesh").interaction == 0
    with pytest.raises(ValueError):
        backend.make_desc(s, keep, interaction_type="LJ")
    assert _lib.OXH_INTERACTION == {"DNA2": 0, "DNA2_nomesh": 0, "RNA2": 1, "RNA": 2}


def test_driver_reads_rna_inputs():
    from oxdna_amd import driver
    base = {"sim_type": "MD", "T": "20C", "dt": "0.005", "seed": "3"}
    o = driver.batch_options(dict(base, interaction_type="RNA2"))
    assert o["interaction_type"] == "RNA2" and o["salt_concentration"] == 1.0   # RNA2's default salt
    o = driver.batch_options(dict(base, interaction_type="RNA", salt_concentration="0.3"))
    assert o["interaction_type"] == "RNA" and o["salt_concentration"] == 0.3
    with pytest.raises(ValueError):
        driver.batch_options(dict(base, interaction_type="RNA2", mismatch_repulsion="true"))
    with pytest.raises(ValueError):
        driver.batch_options(dict(base, interaction_type="RNA2", use_average_seq="false", seq_dep_file_RNA="x"))
    with pytest.raises(KeyError):
        driver.batch_options(dict(base, interaction_type="DNA2"))   # DNA2 needs salt_concentration
    with pytest.raises(ValueError):
        driver.batch_options(dict(base, interaction_type="RNANM"))


def test_uracil_reads_as_thymine():
    from oxdna_amd import io
    assert io.decode_base("U") == io.decode_base("T") == 3
    assert io.decode_base("u") == 3


def test_rna_duplex_geometry():
    """Generated oxRNA duplexes: FENE bonds inside their range, every
    strand-1 nucleotide facing its complementary partner, and the helix
    built from the fitted step (rise and twist per base pair)."""
    from oxdna_amd import systems
    rng = np.random.default_rng(0)
    b = systems._Builder()
    seq = rng.integers(0, 4, 30)
    systems.add_rna_duplex(b, seq, (10.0, 10.0, 10.0), systems.random_rotation(rng))
    s = b.build([20.0] * 3, "r")
    a2 = np.cross(s.a3, s.a1)
    back = s.pos - 0.4 * s.a1 + 0.2 * s.a3
    for i in range(s.N):
        if s.n3[i] >= 0:
            d = np.linalg.norm(back[s.n3[i]] - back[i])
            assert abs(d - 0.761070781051) < 0.25
    assert np.allclose(np.linalg.norm(s.a1, axis=1), 1.0) and np.allclose(np.linalg.norm(s.a3, axis=1), 1.0)
    assert np.allclose(np.einsum("ij,ij->i", s.a1, s.a3), 0.0, atol=1e-12)
    bp = 30
    # strand 1 rows are 3' first: row m holds base pair bp - 1 - m, paired with row bp + (bp - 1 - m)
    for m in range(bp):
        k = bp - 1 - m
        i, j = m, bp + k
        assert s.btype[i] + s.btype[j] == 3
        base_i, base_j = s.pos[i] + 0.4 * s.a1[i], s.pos[j] + 0.4 * s.a1[j]
        assert np.linalg.norm(base_j - base_i) < 0.75   # within HB range (RCHIGH 0.7838)
        assert np.dot(s.a1[i], s.a1[j]) < -0.8
    del a2


def _fixture(name):
    z = dict(np.load(os.path.join(RNA_DIR, name + ".npz")))
    # oxRNA without Debye-Hueckel splits its energy into seven terms
    if len(z["E8"]) == 7:
        z["E8"] = np.append(z["E8"], 0.0)
    return z


def test_reference_binary_reproduces_the_published_rna_split():
    """The fixture of the reference's oxRNA2 force-field test
    (test/RNA/FORCE_FIELD) equals the split the reference publishes for it
    (AVG_SEQ/reference.dat, 6 decimals)."""
    z = _fixture("rna_ff16")
    got = z["E8"] / len(z["pos"])
    assert np.all(np.abs(got - z["E8_published"]) <= 5.01e-7)


@pytest.mark.parametrize("pair", [("rna_ff16", "rna_ff16_nohb"), ("rna_duplex40_37C_0.5M", "rna_duplex40_nohb"),
                                  ("rna_contacts_45C", "rna_contacts_nohb"), ("rna_nicked24_37C", "rna_nicked24_nohb")])
def test_nohb_twins_differ_in_hb_only(pair):
    """Replacing every base type by A leaves all oxRNA2 terms but HB as
    they were (average-sequence stacking, cross and coaxial stacking and
    Debye-Hueckel do not depend on the base), which the twins' reference
    outputs confirm bit for bit."""
    a, b = _fixture(pair[0]), _fixture(pair[1])
    assert np.array_equal(a["pos"], b["pos"]) and np.array_equal(a["a1"], b["a1"])
    cols = [0, 1, 2, 3, 5, 6, 7]
    assert np.array_equal(a["E8"][cols], b["E8"][cols])
    assert a["E8"][4] < 0 and b["E8"][4] == 0 and int(b["hb_count"]) == 0


def test_rna_fixtures_cover_every_term():
    """Across the oxRNA fixtures every one of the eight terms is nonzero
    somewhere, and in a state without HB (where the bar is FP64 rounding)."""
    names = sorted(f[:-4] for f in os.listdir(RNA_DIR) if f.endswith(".npz"))
    nz = np.zeros(8, bool)
    nz_exact = np.zeros(8, bool)
    for n in names:
        z = _fixture(n)
        e = np.abs(z["E8"]) > 1e-9
        nz |= e
        if z["E8"][4] == 0:
            nz_exact |= e
    assert nz.all()
    assert nz_exact[[0, 1, 2, 3, 5, 6, 7]].all()
