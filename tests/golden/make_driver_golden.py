"""Writes tests/golden/driver_nve/ with the reference binary: an input file,
topology and configuration, and the reference's own output files for them.

Run in the build container (needs oracle/_ref/oxDNA, built from the
reference's sources by `make -f oracle/ref.mk`):

    python tests/golden/make_driver_golden.py

The run: a 20-bp duplex with seeded thermal velocities and angular momenta,
300 NVE steps (thermostat = no) of the reference's CPU backend with
interaction_type = DNA2_nomesh, energy every 50 steps, a configuration every
100 steps, the last configuration at the end, in ASCII and in the
reference's binary format (lastconf_file_bin), and a data_output file of
observables (step, energy split, kinetic energy, hydrogen-bond count,
potential energy) every 50 steps.  The fixture keeps the input
files and the reference's energy file, trajectory and last configurations;
the driver test runs the same input through oxdna_amd.driver and compares.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oxdna_amd import io, systems  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "oxDNA")
OUT = os.path.join(HERE, "driver_nve")

INPUT = {
    "backend": "CPU", "sim_type": "MD", "interaction_type": "DNA2_nomesh", "steps": "300", "seed": "7",
    "T": "20C", "salt_concentration": "0.5", "dt": "0.003", "verlet_skin": "0.05", "thermostat": "no",
    "refresh_vel": "0", "restart_step_counter": "1", "topology": "t.top", "conf_file": "c.dat",
    "trajectory_file": "trajectory.dat", "energy_file": "energy.dat", "lastconf_file": "last_conf.dat",
    "print_energy_every": "50", "print_conf_interval": "100", "no_stdout_energy": "1", "time_scale": "linear",
    "box_type": "orthogonal", "log_file": "log.dat", "lastconf_file_bin": "last_conf.bin",
}


# observables of the reference (src/Observables/Step.cpp, PotentialEnergy.cpp,
# KineticEnergy.cpp, HBList.cpp) in one data_output file
OBSERVABLES = """data_output_1 = {
	name = obs.dat
	print_every = 50
	col_1 = {
		type = step
	}
	col_2 = {
		type = potential_energy
		split = true
	}
	col_3 = {
		type = kinetic_energy
		precision = 8
	}
	col_4 = {
		type = hb_list
		only_count = true
	}
	col_5 = {
		type = potential_energy
		general_format = true
		precision = 10
	}
}
"""


# external forces of every supported type, acting from step 500 on (the
# configuration's step counter is kept), on a duplex whose second strand
# starts one box length away in x, so that the run re-boxes it (fix_diffusion
# every 100 steps) and prints absolute positions afterwards
EXT_FORCES = """{
type = string
particle = 0-3
F0 = 0.3
rate = 2e-4
dir = 1, 0.5, 0
}
{
type = trap
particle = 10
stiff = 1.0
rate = 1e-4
dir = 0, 0, 1
pos0 = 10, 10, 10
}
{
type = mutual_trap
particle = 5
ref_particle = 34
stiff = 1.5
r0 = 1.2
PBC = 1
rate = 1e-5
stiff_rate = 1e-4
}
{
type = constant_trap
particle = 7
ref_particle = 30
stiff = 0.2
r0 = 3
}
{
type = repulsion_plane
particle = -1
stiff = 2
dir = 0, 0, -1
position = 13
v = -1e-3
end_position = 12.6
}
{
type = sphere
particle = all
stiff = 1.5
r0 = 3.5
rate = 1e-4
center = 10, 10, 10
}
"""

EXT_INPUT = dict(INPUT, restart_step_counter="0", external_forces="1", external_forces_file="ext.txt",
                 fix_diffusion="1", fix_diffusion_every="100", print_conf_interval="100")


def write_case(out, s, inp, files=(), observables=""):
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        io.write_topology(os.path.join(d, "t.top"), s)
        io.write_conf(os.path.join(d, "c.dat"), s, step=int(inp.get("_step", 0)))
        for name, text in files:
            with open(os.path.join(d, name), "w") as f:
                f.write(text)
        with open(os.path.join(d, "input"), "w") as f:
            for k, v in inp.items():
                if not k.startswith("_"):
                    f.write(f"{k} = {v}\n")
            f.write(observables)
        r = subprocess.run([REF_BIN, "input"], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference failed: {r.stdout[-2000:]} {r.stderr[-2000:]}")
        os.makedirs(out, exist_ok=True)
        pairs = [("input", "input"), ("t.top", "t.top"), ("c.dat", "c.dat"), ("energy.dat", "energy_ref.dat"),
                 ("trajectory.dat", "trajectory_ref.dat"), ("last_conf.dat", "last_conf_ref.dat"),
                 ("last_conf.bin", "last_conf_ref.bin")] + [(n, n) for n, _ in files]
        if observables:
            pairs.append(("obs.dat", "obs_ref.dat"))
        for src_, dst in pairs:
            shutil.copy(os.path.join(d, src_), os.path.join(out, dst))
    print("wrote", out)


def main():
    ext_case()
    s = systems.duplex_20bp(3)
    rng = np.random.default_rng(11)
    s.vel = rng.normal(scale=np.sqrt(s.T), size=(s.N, 3))
    s.vel -= s.vel.mean(axis=0)
    s.L = rng.normal(scale=np.sqrt(s.T), size=(s.N, 3))
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        io.write_topology(os.path.join(d, "t.top"), s)
        io.write_conf(os.path.join(d, "c.dat"), s)
        with open(os.path.join(d, "input"), "w") as f:
            for k, v in INPUT.items():
                f.write(f"{k} = {v}\n")
            f.write(OBSERVABLES)
        r = subprocess.run([REF_BIN, "input"], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference failed: {r.stdout[-2000:]} {r.stderr[-2000:]}")
        os.makedirs(OUT, exist_ok=True)
        for src, dst in (("input", "input"), ("t.top", "t.top"), ("c.dat", "c.dat"), ("energy.dat", "energy_ref.dat"),
                         ("trajectory.dat", "trajectory_ref.dat"), ("last_conf.dat", "last_conf_ref.dat"),
                         ("last_conf.bin", "last_conf_ref.bin"), ("obs.dat", "obs_ref.dat")):
            shutil.copy(os.path.join(d, src), os.path.join(OUT, dst))
    print("wrote", OUT)


def ext_case():
    s = systems.duplex_20bp(3)
    rng = np.random.default_rng(12)
    s.vel = rng.normal(scale=np.sqrt(s.T), size=(s.N, 3))
    s.vel -= s.vel.mean(axis=0)
    s.L = rng.normal(scale=np.sqrt(s.T), size=(s.N, 3))
    s.pos[s.strand == 1] += np.array([s.box[0], 0.0, 0.0])
    write_case(os.path.join(HERE, "driver_ext"), s, dict(EXT_INPUT, _step=500), files=[("ext.txt", EXT_FORCES)])


if __name__ == "__main__":
    main()
