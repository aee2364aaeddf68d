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
