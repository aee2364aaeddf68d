"""Reading external_forces_file (oxdna_amd.io.read_external_forces): the
reference's plain and JSON formats (src/Forces/ForceFactory.cpp:113-200)
and its particle selections (Utils::get_particles_from_string,
src/Utilities/Utils.cpp:409-500), on the fixtures' own force files."""
import json
import os

import numpy as np
import pytest

from _common import GOLDEN
from oxdna_amd import io, systems


def _fixture(name):
    z = np.load(os.path.join(GOLDEN, "external", name + ".npz"))
    s = systems.System(pos=z["pos"], a1=z["a1"], a3=z["a3"], btype=z["btype"], n3=z["n3"], n5=z["n5"],
                       strand=z["strand"], box=z["box"])
    return s, str(z["forces_file"])


def test_plain_format_expands_particle_lists(tmp_path):
    s, text = _fixture("ext_duplex_step0")
    p = tmp_path / "f.txt"
    p.write_text(text)
    recs = io.read_external_forces(str(p), s)
    by_type = {}
    for r in recs:
        by_type.setdefault(r["type"], []).append(r["particle"])
    assert by_type["string"] == [0, 1, 2, 3, 39]
    assert by_type["trap"] == [10, 12, s.N - 1]
    assert by_type["mutual_trap"] == [5] and by_type["constant_trap"] == [7]
    assert by_type["repulsion_plane"] == list(range(s.N)) and by_type["sphere"] == list(range(s.N))
    plane = [r for r in recs if r["type"] == "repulsion_plane"][0]
    assert plane["end_position"] == 12.0 and plane["v"] == -1e-3 and plane["dir"] == (0.0, 0.0, -1.0)
    centre = [r for r in recs if r.get("dir_as_centre")]
    assert len(centre) == 1 and centre[0]["particle"] == 39


def test_ranges_walk_along_the_strand():
    """a-b follows the 5' neighbours from a until b (the topology's n5)."""
    s = systems.hairpin_set(3, n=4)[1]
    assert io.particles_from_string("3-6", s) == [3, 4, 5, 6]
    # against the direction of n5 the walk reaches the strand's end first
    with pytest.raises(ValueError):
        io.particles_from_string("6-3", s)
    assert io.particles_from_string("last, 0", s) == [s.N - 1, 0]


def test_json_format_matches_plain(tmp_path):
    s, text = _fixture("ext_hairpin_pbc")
    p = tmp_path / "f.txt"
    p.write_text(text)
    plain = io.read_external_forces(str(p), s)
    blocks = io._force_blocks(text)
    j = tmp_path / "f.json"
    j.write_text(json.dumps(blocks))
    assert io.read_external_forces(str(j), s, as_json=True) == plain


def test_unsupported_type_raises(tmp_path):
    s = systems.duplex_20bp(0)
    p = tmp_path / "f.txt"
    p.write_text("{\ntype = hard_wall\nparticle = 0\n}\n")
    with pytest.raises(ValueError, match="hard_wall"):
        io.read_external_forces(str(p), s)
    p.write_text("{\ntype = twist\nparticle = 0\nstiff = 1\nrate = 0\n}\n")
    with pytest.raises(ValueError, match="missing key base"):
        io.read_external_forces(str(p), s)


def _read(tmp_path, text, s=None):
    s = s or systems.duplex_20bp(0)
    p = tmp_path / "f.txt"
    p.write_text(text)
    return io.read_external_forces(str(p), s)


def test_keys_and_defaults_follow_the_reference_classes(tmp_path):
    """Each class's init: its defaults (LJWall.cpp:17-25, RepulsiveEllipsoid.cpp:
    15-19, RepulsiveKeplerPoinsot.cpp:34-44, MutualTrap.cpp:17-34,
    RepulsionPlane.h:41-42) and its key handling."""
    r = _read(tmp_path, "{\ntype = LJ_wall\nparticle = 1\nposition = 2\ndir = 0,0,1\n}\n")[0]
    assert (r["stiff"], r["sigma"], r["n"], r["cutoff"]) == (1.0, 1.0, 6, 1e6)
    r = _read(tmp_path, "{\ntype = LJ_cone\nparticle = 1\nalpha = 0.5\nn = 8\nonly_repulsive = true\n"
                        "dir = 0,0,1\npos0 = 0,0,0\n}\n")[0]
    assert r["cutoff"] == 2.0 ** (1.0 / 8) and r["n"] == 8
    with pytest.raises(ValueError, match="even"):
        _read(tmp_path, "{\ntype = LJ_wall\nparticle = 1\nposition = 2\nn = 5\ndir = 0,0,1\n}\n")
    r = _read(tmp_path, "{\ntype = ellipsoid\nparticle = 2\nstiff = 1\nr_2 = 1,2,3\n}\n")[0]
    assert r["r_1"] == (1e-6, 1e-6, 1e-6) and r["center"] == (0.0, 0.0, 0.0)
    r = _read(tmp_path, "{\ntype = repulsive_kepler_poinsot\nparticle = 0\nstiff = 1\nkappa = -3\n}\n")[0]
    assert (r["apex"], r["base"], r["base_radius"], r["kappa"], r["rate"]) == (1.2, 0.7, 0.45, 25.0, 0.0)
    with pytest.raises(ValueError, match="apex > base"):
        _read(tmp_path, "{\ntype = repulsive_kepler_poinsot\nparticle = 0\nstiff = 1\napex = 0.5\n}\n")
    r = _read(tmp_path, "{\ntype = mutual_trap\nparticle = 0\nref_particle = 3\nstiff = 1\nr0 = 1\n}\n")[0]
    assert r["PBC"] is False and r["rate"] == 0.0 and r["stiff_rate"] == 0.0
    r = _read(tmp_path, "{\ntype = repulsion_plane\nparticle = 0\nstiff = 1\nposition = 1\ndir = 1,0,0\n}\n")[0]
    assert r["end_position"] == 1e6 and r["v"] == 0.0
    r = _read(tmp_path, "{\ntype = twist\nparticle = 0\nstiff = 1\nrate = 0\nbase = 0.2\naxis = 0,0,1\n"
                        "center = 0,0,0\npos0 = 1,0,0\n}\n")[0]
    assert r["F0"] == 0.2 and r["mask"] == (0.0, 0.0, 0.0)
    r = _read(tmp_path, "{\ntype = repulsive_sphere_moving\nparticle = 0\nstiff = 1\nr0 = 1\ncenter = 1,2,3\n"
                        "move_steps = 500\n}\n")[0]
    assert r["pos0"] == (1.0, 2.0, 3.0) and r["target"] == (0.0, 0.0, 0.0) and r["steps"] == 500
    r = _read(tmp_path, "{\ntype = lowdim_trap\nparticle = 0-2\nstiff = 1\nrate = 0\npos0 = 0,0,0\n"
                        "dir = 1,0,0\nvisibility = 1, 0, 1\n}\n")
    assert [x["particle"] for x in r] == [0, 1, 2] and r[0]["visibility"] == (1, 0, 1)


def test_reference_key_quirks_are_kept(tmp_path):
    """Where a class reads a key into another field, the value is the one
    the reference runs with: RepulsiveSphereSmooth takes smooth and alpha
    from r_ext (RepulsiveSphereSmooth.cpp:30-32); YukawaSphere reads WCA_n
    into the WCA sigma and keeps n = 6 (YukawaSphere.cpp:27-29)."""
    r = _read(tmp_path, "{\ntype = sphere_smooth\nparticle = 0\nstiff = 1\nr0 = 1\nr_ext = 4\n}\n")[0]
    assert r["smooth"] == r["alpha"] == r["r_ext"] == 4.0
    with pytest.raises(ValueError, match="r_ext"):
        _read(tmp_path, "{\ntype = sphere_smooth\nparticle = 0\nstiff = 1\nr0 = 1\n}\n")
    r = _read(tmp_path, "{\ntype = yukawa_sphere\nparticle = 0\nradius = 5\ndebye_length = 2\ndebye_A = 1\n"
                        "WCA_sigma = 1.5\nWCA_n = 3\n}\n")[0]
    assert r["sigma"] == 3.0 and r["n"] == 6 and r["cutoff"] == 8.0 and r["epsilon"] == 1.0


def test_com_and_moving_plane_lists(tmp_path):
    s = systems.duplex_20bp(0)
    recs = _read(tmp_path, "{\ntype = com\ncom_list = 0-3\nref_list = 30-31, 35\nstiff = 1\nr0 = 2\n}\n", s)
    assert [r["particle"] for r in recs] == [0, 1, 2, 3]
    assert recs[0]["com_list"] == (0, 1, 2, 3) and recs[0]["ref_list"] == (30, 31, 35) and recs[0]["rate"] == 0.0
    recs = _read(tmp_path, "{\ntype = repulsion_plane_moving\nparticle = 20\nref_particle = 7, 5, 6\nstiff = 1\n"
                           "dir = 0,0,1\n}\n", s)
    assert recs[0]["ref_particle"] == 5 and recs[0]["ref_last"] == 7
    with pytest.raises(ValueError, match="contiguous"):
        _read(tmp_path, "{\ntype = repulsion_plane_moving\nparticle = 20\nref_particle = 5, 7\nstiff = 1\n"
                        "dir = 0,0,1\n}\n", s)
    with pytest.raises(ValueError, match="interpolated"):
        _read(tmp_path, "{\ntype = generic_central_force\nparticle = 0\nforce_type = interpolated\n"
                        "center = 0,0,0\n}\n", s)


@pytest.mark.parametrize("name", sorted(os.path.basename(f)[:-4] for f in
                                        __import__("glob").glob(os.path.join(GOLDEN, "external", "*.npz"))))
def test_every_fixture_force_file_parses(name):
    s, text = _fixture(name)
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "f.txt")
        with open(p, "w") as f:
            f.write(text)
        recs = io.read_external_forces(p, s)
    assert recs and all(0 <= r["particle"] < s.N for r in recs)
