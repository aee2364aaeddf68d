"""Writes tests/golden/order_params/*.json with the reference binary: the
order_parameters observable (src/Observables/OrderParameterValues.cpp) of
fixture states, for order-parameter files built here.

Run in the build container (needs /root/reference and oracle/_ref/oxDNA,
built by `make -f oracle/ref.mk`):

    python tests/golden/make_op_golden.py

For each case the script takes a force fixture's state and writes an
order_parameters_file with:
  - bond parameters: the complementary, non-bonded pairs whose base sites
    are within 1.2 (hydrogen-bond candidates), and a pair list in the
    range syntax "a1-a2, b1-b2";
  - mindistance parameters over cross-strand and same-strand pairs:
    base-base distances, centre-centre distances (use_COM = true), and a
    copy with interfaces.
It runs the strict reference binary (backend = CPU, DNA2_nomesh, steps =
0) with the observable printed at step 0 and stores the op file and the
printed line.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from _common import load_golden  # noqa: E402
from oxdna_amd import io  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "oxDNA")
OUT = os.path.join(HERE, "order_params")
CASES = ["duplex20_md_20C", "nicked_md_30C", "hairpin_md_50C", "bundle_md_30C", "hot_md_cxst_95C", "seqdep_md_25C"]


def op_file(s) -> str:
    base = s.pos + 0.4 * s.a1
    N = s.N
    near = []
    for i in range(N):
        for j in range(i + 1, N):
            if s.btype[i] + s.btype[j] != 3 or s.n3[i] == j or s.n5[i] == j:
                continue
            d = base[j] - base[i]
            if s.strand[i] != s.strand[j]:
                d -= s.box * np.rint(d / s.box)
            if d @ d < 1.2 ** 2:
                near.append((i, j))
    blocks = []
    lines = [f"pair{k + 1} = {i}, {j}" for k, (i, j) in enumerate(near[:40])] or ["pair1 = 0, 1"]
    blocks.append("order_parameter = bond\nname = near_pairs\n" + "\n".join(lines))
    h = N // 2
    blocks.append(f"order_parameter = bond\nname = range_pairs\npair1 = 0-3, {h}-{h + 3}\npair2 = 5, {N - 6}")
    md = [(0, N - 1), (0, 1), (0, h), (2, N - 3)] + near[:6]
    md_lines = "\n".join(f"pair{k + 1} = {i}, {j}" for k, (i, j) in enumerate(md))
    blocks.append("order_parameter = mindistance\nname = md_base\n" + md_lines)
    blocks.append("order_parameter = mindistance\nname = md_com\nuse_COM = true\n" + md_lines)
    blocks.append("order_parameter = mindistance\nname = md_far\ninterfaces = 0.5, 1.0, 2.0\n"
                  f"pair1 = 0, {h}\npair2 = 1, {N - 2}")
    return "".join("{\n" + b + "\n}\n" for b in blocks)


def run(name: str) -> dict:
    s, opts, _ = load_golden(name)
    text = op_file(s)
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        io.write_topology(os.path.join(d, "t.top"), s)
        io.write_conf(os.path.join(d, "c.dat"), s)
        with open(os.path.join(d, "op.txt"), "w") as f:
            f.write(text)
        inp = {"backend": "CPU", "sim_type": "MD", "interaction_type": "DNA2_nomesh", "steps": "0", "seed": "1",
               "T": f"{s.T}", "salt_concentration": repr(s.salt), "dt": "0.003", "verlet_skin": "0.05",
               "thermostat": "no", "refresh_vel": "1", "topology": "t.top", "conf_file": "c.dat",
               "trajectory_file": "traj.dat", "energy_file": "energy.dat", "print_energy_every": "1000",
               "print_conf_interval": "100000", "no_stdout_energy": "1", "restart_step_counter": "1",
               "box_type": "orthogonal", "time_scale": "linear", "log_file": "log.dat"}
        if "seq_stck" in opts:
            inp["use_average_seq"] = "false"
            inp["seq_dep_file"] = "/root/reference/oxDNA2_sequence_dependent_parameters.txt"
        with open(os.path.join(d, "input"), "w") as f:
            for k, v in inp.items():
                f.write(f"{k} = {v}\n")
            f.write("data_output_1 = {\n name = op.dat\n print_every = 1\n col_1 = {\n type = order_parameters\n"
                    " order_parameters_file = op.txt\n }\n}\n")
        r = subprocess.run([REF_BIN, "input"], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference failed: {r.stdout[-2000:]} {r.stderr[-2000:]}")
        line = open(os.path.join(d, "op.dat")).read().splitlines()[0]
    return {"case": name, "op_file": text, "expected": line}


def main():
    os.makedirs(OUT, exist_ok=True)
    for name in CASES:
        res = run(name)
        with open(os.path.join(OUT, name + ".json"), "w") as f:
            json.dump(res, f, indent=1)
        print(name, repr(res["expected"]))


if __name__ == "__main__":
    main()
