"""Writes tests/golden/ensemble/duplex20.npz with the reference binary:
equilibrium time series of the 20-bp duplex (the reference's CPU-runnable
case: random ACGT sequence, cubic box of 20, 0.5 M, dt 0.003, Langevin with
diff_coeff 2.5) at 20 C and at 50 C.

Run in the build container (needs oracle/_ref/oxDNA, the strict-IEEE FP64
build of the reference's sources that pins the oracle, built by
`make -f oracle/ref.mk`):

    python tests/golden/make_ensemble_golden.py

For each temperature, RUNS independent runs (seeds SEED0 + k) of STEPS MD
steps each, all from the same ideal helix with fresh thermal velocities,
with backend = CPU and interaction_type = DNA2_nomesh print, every
EVERY steps, the energy file (potential and kinetic energy per nucleotide,
src/Observables/TotalEnergy.cpp) and the hb_list observable with
only_count = true (pairs with hydrogen-bonding energy below HB_CUTOFF = -0.1,
src/Observables/HBList.cpp:64-164, src/defs.h:62).  The fixture keeps those
series (float32 energies, int8 counts) and the run parameters; the GPU test
runs the same protocol and compares the means within two standard errors.
The duplex's slow modes (fraying ends) decorrelate over ~1e5 steps or more,
so the statistics use the means of whole independent runs.

The strict binary is the reference here: at 50 C the release build
(-O3 -ffast-math) gave U/N lower by 0.005 (about two standard errors of 144
runs) than the strict build and this library, which agree.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oxdna_amd import io, systems  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "oxDNA")
OUT = os.path.join(HERE, "ensemble", "duplex20.npz")

SYSTEM_SEED = 0
TEMPS_C = (20.0, 50.0)
RUNS = 192
SEED0 = 2001
STEPS = 300_000
EVERY = 1000

INPUT = """backend = CPU
sim_type = MD
interaction_type = DNA2_nomesh
steps = {steps}
seed = {seed}
T = {tc}C
salt_concentration = 0.5
dt = 0.003
verlet_skin = 0.05
thermostat = langevin
diff_coeff = 2.5
refresh_vel = 1
restart_step_counter = 1
topology = t.top
conf_file = c.dat
trajectory_file = trajectory.dat
energy_file = energy.dat
lastconf_file = last_conf.dat
print_energy_every = {every}
print_conf_interval = 1000000000
no_stdout_energy = 1
time_scale = linear
log_file = log.dat
data_output_1 = {{
	name = hb.dat
	print_every = {every}
	col_1 = {{
		type = hb_list
		only_count = true
	}}
}}
"""


def main():
    s = systems.duplex_20bp(SYSTEM_SEED)
    jobs = []
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        for ti, tc in enumerate(TEMPS_C):
            for k in range(RUNS):
                wd = os.path.join(d, f"t{ti}_r{k}")
                os.makedirs(wd)
                io.write_topology(os.path.join(wd, "t.top"), s)
                io.write_conf(os.path.join(wd, "c.dat"), s)
                with open(os.path.join(wd, "input"), "w") as f:
                    f.write(INPUT.format(steps=STEPS, seed=SEED0 + k, tc=tc, every=EVERY))
                jobs.append((ti, k, wd))
        procs = []
        for ti, k, wd in jobs:
            while sum(p.poll() is None for _, _, _, p in procs) >= os.cpu_count():
                procs[[p.poll() is None for _, _, _, p in procs].index(True)][3].wait()
            procs.append((ti, k, wd, subprocess.Popen([REF_BIN, "input"], cwd=wd, stdout=subprocess.DEVNULL,
                                                      stderr=subprocess.DEVNULL)))
        n = STEPS // EVERY + 1
        U = np.zeros((len(TEMPS_C), RUNS, n), np.float32)
        K = np.zeros_like(U)
        hb = np.zeros((len(TEMPS_C), RUNS, n), np.int8)
        for ti, k, wd, p in procs:
            if p.wait() != 0:
                raise RuntimeError(f"reference run {wd} failed")
            e = np.loadtxt(os.path.join(wd, "energy.dat"))
            c = np.loadtxt(os.path.join(wd, "hb.dat"))
            U[ti, k], K[ti, k], hb[ti, k] = e[:, 1], e[:, 2], c
    np.savez_compressed(OUT, U=U, K=K, hb=hb, temps_c=np.array(TEMPS_C), every=EVERY, steps=STEPS, runs=RUNS,
                        seed0=SEED0, system_seed=SYSTEM_SEED)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
