"""FFS order parameters on the host side: the op-file parser against the
reference's rules (src/Utilities/OrderParameters.cpp:301-598) and the
mindistance formula against the reference binary's printed values.

The fixtures (tests/golden/order_params/, written by
tests/golden/make_op_golden.py) hold an order_parameters_file and the line
the reference's order_parameters observable printed for a fixture state:
bond counts, then mindistance values (6 significant digits).  Here the
distances are recomputed in FP64 with numpy from the state (base site =
centre + 0.4 a1, the plain difference for two particles of one strand and
the nearest image otherwise), which pins the formula the GPU kernel
follows; the bond counts need the hydrogen-bonding energy and are checked
on the GPU (tests/test_gpu_order_params.py)."""
import glob
import json
import os

import numpy as np
import pytest

from _common import GOLDEN, load_golden

CASES = sorted(os.path.basename(f)[:-5] for f in glob.glob(os.path.join(GOLDEN, "order_params", "*.json")))


def load(name, tmp_path):
    from oxdna_amd import io
    c = json.load(open(os.path.join(GOLDEN, "order_params", name + ".json")))
    s, opts, _ = load_golden(c["case"])
    p = tmp_path / "op.txt"
    p.write_text(c["op_file"])
    return s, opts, io.read_order_parameters(str(p), s.N), c["expected"].split()


def test_cases_exist():
    assert len(CASES) >= 5


def test_range_syntax_pairs_a_range_with_a_reversed_range(tmp_path):
    from oxdna_amd import io
    p = tmp_path / "op.txt"
    p.write_text("{\norder_parameter = bond\nname = x\npair1 = 0-3, 20-23\npair2 = 5, 2\npair3 = 2, 5\n}\n"
                 "{\norder_parameter = mindistance\nname = y\nuse_COM = true\ninterfaces = 0.5, 1\npair1 = 1, 0\n}\n")
    ps = io.read_order_parameters(str(p), 40)
    assert ps[0]["type"] == "bond" and ps[0]["pairs"] == [(0, 23), (1, 22), (2, 5), (2, 21), (3, 20)]
    assert ps[1]["type"] == "mindistance" and ps[1]["pairs"] == [(0, 1)] and ps[1]["use_COM"]
    assert ps[1]["interfaces"] == [0.5, 1.0]


@pytest.mark.parametrize("text", [
    "{\norder_parameter = twist\nname = t\npair1 = 0, 1\n}\n",                 # not supported
    "{\norder_parameter = bond\nname = b\npair1 = 0-3, 20\n}\n",                # one side a range
    "{\norder_parameter = bond\nname = b\npair1 = 0-3, 20-22\n}\n",             # unequal ranges
    "{\norder_parameter = bond\nname = b\npair1 = 0, 40\n}\n",                  # out of range
    "{\norder_parameter = bond\npair1 = 0, 1\n}\n",                             # no name
    "{\norder_parameter = bond\nname = b\npair1 = 0, 1\n",                      # unbalanced
])
def test_malformed_files_raise(tmp_path, text):
    from oxdna_amd import io
    p = tmp_path / "op.txt"
    p.write_text(text)
    with pytest.raises((ValueError, KeyError)):
        io.read_order_parameters(str(p), 40)


@pytest.mark.parametrize("name", CASES)
def test_mindistance_formula_reproduces_the_reference(tmp_path, name):
    s, _, params, want = load(name, tmp_path)
    n_bond = sum(p["type"] == "bond" for p in params)
    dists = []
    for p in params:
        if p["type"] != "mindistance":
            continue
        best = None
        for i, j in p["pairs"]:
            d = s.pos[j] - s.pos[i]
            if s.strand[i] != s.strand[j]:
                d = d - s.box * np.rint(d / s.box)
            if not p["use_COM"]:
                d = d + 0.4 * s.a1[j] - 0.4 * s.a1[i]
            best = d @ d if best is None else min(best, d @ d)
        dists.append(np.sqrt(best))
    assert len(want) == n_bond + len(dists)
    for got, w in zip(dists, want[n_bond:]):
        assert f"{got:g}" == w
