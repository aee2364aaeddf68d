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
    p.write_text("{\ntype = twist\nparticle = 0\n}\n")
    with pytest.raises(ValueError, match="twist"):
        io.read_external_forces(str(p), s)
