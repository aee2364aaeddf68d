"""CPU tests of the host driver (oxdna_amd/driver.py) against the files the
reference binary wrote for the same input (tests/golden/driver_nve/,
written by tests/golden/make_driver_golden.py): the input is understood,
the output schedule is the reference's, and energy lines are formatted as
the reference formats them."""
import os

import numpy as np

from _common import GOLDEN
from oxdna_amd import driver, io

D = os.path.join(GOLDEN, "driver_nve")


def _confs(path):
    """[(step, rows)] of a trajectory or configuration file."""
    out = []
    with open(path) as f:
        lines = f.read().splitlines()
    i = 0
    while i < len(lines):
        assert lines[i].startswith("t = ")
        step = int(lines[i].split("=")[1])
        j = i + 3
        while j < len(lines) and not lines[j].startswith("t = "):
            j += 1
        out.append((step, np.array([[float(x) for x in l.split()] for l in lines[i + 3:j]])))
        i = j
    return out


def test_driver_reads_reference_input():
    inp = io.read_input(os.path.join(D, "input"))
    opts = driver.batch_options(inp)
    assert opts["thermostat"] == "no" and opts["dt"] == 0.003 and opts["salt_concentration"] == 0.5
    assert abs(opts["T"] - io.temperature("20C")) < 1e-15


def test_driver_refuses_what_it_does_not_simulate():
    """Inputs the driver cannot run as the reference would are refused,
    never run without the missing part."""
    import pytest
    base = io.read_input(os.path.join(D, "input"))
    for key, val in (("sim_type", "MC"), ("interaction_type", "DNA")):
        with pytest.raises(ValueError, match=key):
            driver.batch_options(dict(base, **{key: val}))
    # sequence dependence without its parameter file
    with pytest.raises(ValueError, match="seq_dep_file"):
        driver.batch_options(dict(base, use_average_seq="false"))
    assert driver.batch_options(dict(base, external_forces="false"))["dt"] == 0.003


def test_driver_reads_seq_dep_file(tmp_path):
    """use_average_seq = false reads the file's well depths in FP32, as the
    reference does; they are the ones the fixture seqdep_md_25C pins."""
    from _common import load_golden
    p = tmp_path / "seq.txt"
    keys = ["STCK_FACT_EPS = 0.18"]
    names = "AGCT"
    _, want, _ = load_golden("seqdep_md_25C")
    for i in range(4):
        for j in range(4):
            keys.append(f"STCK_{names[i]}_{names[j]} = {float(want['seq_stck'][i][j])!r}")
    keys += [f"HYDR_T_A = {float(want['seq_hydr_AT'])!r}", f"HYDR_C_G = {float(want['seq_hydr_GC'])!r}"]
    p.write_text("\n".join(keys) + "\n")
    base = io.read_input(os.path.join(D, "input"))
    got = driver.batch_options(dict(base, use_average_seq="false", seq_dep_file=str(p)))
    assert np.array_equal(np.array(got["seq_stck"]), want["seq_stck"])
    assert got["seq_stck_fact_eps"] == want["seq_stck_fact_eps"]
    assert got["seq_hydr_AT"] == want["seq_hydr_AT"] and got["seq_hydr_GC"] == want["seq_hydr_GC"]


def test_seed_defaults_like_the_reference():
    """A given seed is kept; without one (or with 0) each run draws a new
    one from the clock and /dev/urandom (src/Managers/SimManager.cpp:
    78-90), so replicas launched from one input differ."""
    base = io.read_input(os.path.join(D, "input"))
    assert driver.batch_options(dict(base, seed="12345"))["seed"] == 12345
    no_seed = {k: v for k, v in base.items() if k != "seed"}
    seeds = {driver.batch_options(no_seed)["seed"] for _ in range(8)}
    assert len(seeds) > 1 and 0 not in seeds
    assert driver.batch_options(dict(base, seed="0"))["seed"] != 0


def test_output_schedule_matches_reference_files():
    energy = np.loadtxt(os.path.join(D, "energy_ref.dat"), ndmin=2)
    traj = _confs(os.path.join(D, "trajectory_ref.dat"))
    plan = driver.output_steps(0, 300, 50, 100)
    e_steps = [t for t, e, _ in plan if e]
    c_steps = [t for t, _, c in plan if c]
    assert np.allclose(np.array(e_steps) * 0.003, energy[:, 0])
    assert c_steps == [s for s, _ in traj]
    # a configuration is not printed at the starting step; energies are
    assert driver.output_steps(100, 200, 50, 100)[0] == (100, True, False)


def test_energy_line_format_matches_reference():
    with open(os.path.join(D, "energy_ref.dat")) as f:
        ref = [l.rstrip() for l in f]
    for k, line in enumerate(ref):
        t, U, K, tot = (float(x) for x in line.split())
        mine = driver.energy_line(round(t / 0.003), 0.003, U, K)
        # same columns and widths; the total may differ in its last digit
        # (the reference adds the unrounded values)
        assert mine.split()[:3] == line.split()[:3]
        assert len(mine) == len(line)


def test_binary_configuration_matches_reference():
    """The reference's binary last configuration (lastconf_file_bin, written
    beside the ASCII one at step 300) reads back as the ASCII file's values
    (ASCII prints 15 significant digits), with the step and the per-particle
    energies of its header."""
    st, box, pos, a1, a2, a3, vel, L, seed = io.read_conf_binary(os.path.join(D, "last_conf_ref.bin"), 40)
    st2, box2, pos2, b1, b2, b3, vel2, L2 = io.read_conf(os.path.join(D, "last_conf_ref.dat"), 40)
    assert st == st2 == 300 and np.array_equal(box, box2)
    for x, y in ((pos, pos2), (a1, b1), (a2, b2), (a3, b3), (vel, vel2), (L, L2)):
        assert np.abs(x - y).max() <= 1e-13 * max(1.0, np.abs(y).max())
    E_ascii = [float(x) for x in open(os.path.join(D, "last_conf_ref.dat")).read().splitlines()[2].split("=")[1].split()]
    raw = open(os.path.join(D, "last_conf_ref.bin"), "rb").read()
    E_bin = np.frombuffer(raw, "<f8", 3, 8 + 6 + 24)
    assert np.allclose(E_bin, E_ascii, rtol=1e-13)


def test_binary_configuration_writer_layout():
    """write_conf_binary writes the reference's layout byte for byte: the
    reference's own file, read and written again with its header values,
    differs only in the a2 axis (recomputed as a3 x a1, last-bit
    differences) and reads back exactly."""
    import tempfile
    ref = open(os.path.join(D, "last_conf_ref.bin"), "rb").read()
    st, box, pos, a1, a2, a3, vel, L, seed = io.read_conf_binary(os.path.join(D, "last_conf_ref.bin"), 40)
    E = np.frombuffer(ref, "<f8", 3, 8 + 6 + 24)
    s = io.load_system(os.path.join(D, "t.top"), os.path.join(D, "last_conf_ref.bin"), binary=True)
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "x.bin")
        io.write_conf_binary(out, s, step=st, energies=tuple(E), seed=seed)
        mine = open(out, "rb").read()
        back = io.read_conf_binary(out, 40)
    assert len(mine) == len(ref)
    head, rec = 62, 156
    a2_bytes = set()
    for i in range(40):
        o = head + i * rec + 24 + 12 + 24
        a2_bytes.update(range(o, o + 24))
    assert all(mine[k] == ref[k] for k in range(len(ref)) if k not in a2_bytes)
    for x, y in zip(back[2:8], (pos, a1, np.cross(a3, a1), a3, vel, L)):
        assert np.array_equal(x, y)


def test_data_output_block_parsed_like_the_reference(tmp_path):
    """The reference's data_output block of the golden input: the file name,
    the schedule (every 50 steps, step 0 and the last step included, as the
    reference's obs_ref.dat) and the columns with their formats."""
    inp = io.read_input(os.path.join(D, "input"))
    o = driver.DataOutput(inp["data_output_1"], 0.003, lambda p: str(tmp_path / p))
    try:
        ref = np.loadtxt(os.path.join(D, "obs_ref.dat"), ndmin=2)
        assert o.steps(0, 300) == [int(x) for x in ref[:, 0]]
        assert [c["type"] for c in o.cols] == ["step", "potential_energy", "kinetic_energy", "hb_list", "potential_energy"]
        assert o.cols[1]["split"] and o.cols[2]["fmt"] == "{:10.8f}" and o.cols[4]["fmt"] == "{:.10g}"
        # the reference's line layout: a step column, then the split terms
        # each preceded by a space
        first = open(os.path.join(D, "obs_ref.dat")).readline()
        assert first.startswith(f"{0:12d}  " + f"{ref[0, 1]:10.6f}")
    finally:
        o.close()
