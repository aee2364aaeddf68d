"""FFS order parameters on the GPU (oxh_batch_pair_quantities, the
driver's order_parameters observable), through the C ABI.

  - Per-pair hydrogen-bonding energies of every pair of every force
    fixture, in FP64: their sum equals the reference binary's HB energy
    (the HYDROGEN_BONDING term of its energy split) and the pairs below
    -0.1 number the reference's hb_list count;
  - the order_parameters observable of fixture states equals the line the
    reference binary printed for the same order_parameters_file
    (tests/golden/order_params/): bond counts exactly, mindistance values
    at the 6 printed digits;
  - the "nearly bonded" criterion of the CUDA FFS backend: every pair with
    a negative HB energy (all seven factors active) is nearly bonded;
    bonded neighbours and non-complementary pairs never are.
"""
import glob
import json
import os

import numpy as np
import pytest

from _common import GOLDEN, golden_names, load_golden

pytestmark = pytest.mark.gpu

OP_CASES = sorted(os.path.basename(f)[:-5] for f in glob.glob(os.path.join(GOLDEN, "order_params", "*.json")))


@pytest.fixture(scope="module")
def hip():
    from oxdna_amd import _lib, backend
    assert _lib.device_count() >= 1, "no GPU visible"
    return backend


@pytest.mark.parametrize("name", golden_names())
def test_pair_hb_energies_sum_to_the_reference_term(hip, name):
    s, opts, ref = load_golden(name)
    i, j = np.triu_indices(s.N, 1)
    pairs = np.stack([i, j], axis=1)
    with hip.HIPBatch(s, **opts) as b:
        e = b.pair_quantities("hb_energy", pairs)
        near = b.pair_quantities("near_hb", pairs)
    scale = max(1.0, np.abs(ref["E8"]).sum())
    assert abs(e.sum() - ref["E8"][4]) <= 1e-9 * scale, (e.sum(), ref["E8"][4])
    assert int(np.count_nonzero(e < -0.1)) == ref["hb_count"]
    # nearly bonded: implied by a negative HB energy, excluded for bonded
    # neighbours and non-complementary pairs
    assert np.all(near[e < 0.0] == 1.0)
    bonded = (s.n3[i] == j) | (s.n5[i] == j)
    noncomp = s.btype[i] + s.btype[j] != 3
    assert not np.any(near[bonded | noncomp])


@pytest.mark.parametrize("name", OP_CASES)
def test_order_parameters_observable_matches_reference(hip, tmp_path, name):
    from oxdna_amd import driver
    c = json.load(open(os.path.join(GOLDEN, "order_params", name + ".json")))
    s, opts, _ = load_golden(c["case"])
    (tmp_path / "op.txt").write_text(c["op_file"])
    block = "{\n name = op.dat\n print_every = 1\n col_1 = {\n type = order_parameters\n order_parameters_file = op.txt\n }\n}"
    out = driver.DataOutput(block, 0.003, lambda p: str(tmp_path / p), {}, s.N)
    with hip.HIPBatch(s, **opts) as b:
        line = out.line(0, b, s.N)
    out.close()
    got, want = line.split(), c["expected"].split()
    n_bond = sum(p["type"] == "bond" for p in out.cols[0]["params"])
    assert got[:n_bond] == want[:n_bond]
    for g, w in zip(got[n_bond:], want[n_bond:]):
        assert abs(float(g) - float(w)) <= 1e-5 * abs(float(w)), (g, w)
    assert line.endswith(" ")


def test_pair_quantities_reject_bad_input(hip):
    from oxdna_amd import _lib
    s, opts, _ = load_golden("duplex20_md_20C")
    with hip.HIPBatch(s, **opts) as b:
        with pytest.raises(_lib.OxhError):
            b.pair_quantities("hb_energy", [[0, s.N]])
        assert len(b.pair_quantities("base_distance", np.zeros((0, 2)))) == 0
