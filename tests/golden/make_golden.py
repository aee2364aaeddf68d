"""Writes the golden fixtures of tests/golden/ with the reference binary.

Run in the build container (needs /root/reference and oracle/_ref/oxDNA,
built by `make -f oracle/ref.mk`):

    python tests/golden/make_golden.py

For every case it writes a topology + configuration, runs the reference
CPU backend (`backend = CPU`, `interaction_type = DNA2_nomesh`, `steps = 0`)
and records what the reference computes at step 0:
  - the 8-term energy split at 15 decimals (potential_energy observable,
    split = true, precision = 15; src/Observables/PotentialEnergy.cpp:39-55);
  - per-particle forces and body-frame torques at %.17g through the
    project's ForceDump observable plugin (oracle/ForceDump.cpp), loaded by
    the reference's own PluginManager;
  - the stress tensor of the pressure observable and the base-pair count of
    the hb_list observable (only_count = true, src/Observables/HBList.cpp).
Each case becomes one small .npz (inputs, the input-file options that set
up the interaction, and expected outputs only).  Cases whose .npz already
exists are re-run from the stored inputs, and the stored outputs must come
back unchanged; only fields that are missing are added.

The cases cover every term of the energy split at non-ideal geometry:
bonded and non-bonded excluded volume (compressed helices, tangles packed
inside the excluded-volume ranges, frames of reference MD runs of dense
systems picked for those terms), coaxial stacking (helices continuous
through the periodic boundary, thermal frames), and the interaction
options max_backbone_force[_far], hb_multiplier with custom base types
(|btype| >= 300), debye_huckel_rhigh and sequence-dependent well depths
(use_average_seq = false with the reference's own seq_dep_file).
"""
from __future__ import annotations

import math
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
PLUGIN_DIR = os.path.join(ROOT, "oracle", "_ref")
SEQ_FILE = "/root/reference/oxDNA2_sequence_dependent_parameters.txt"
BASES = "AGCT"


def read_seq_file(path: str) -> dict:
    """Stacking and H-bond well depths of a seq_dep_file, rounded to FP32
    as the reference reads them (getInputFloat into float,
    src/Interactions/DNAInteraction.cpp:328-376): stck[4][4] for A, G, C, T,
    STCK_FACT_EPS, HYDR_A_T and HYDR_G_C."""
    kv = {}
    for line in open(path):
        if "=" in line:
            k, v = line.split("=", 1)
            kv[k.strip()] = float(np.float32(float(v)))
    stck = np.array([[kv[f"STCK_{BASES[i]}_{BASES[j]}"] for j in range(4)] for i in range(4)])
    at = kv.get("HYDR_A_T", kv.get("HYDR_T_A"))
    gc = kv.get("HYDR_G_C", kv.get("HYDR_C_G"))
    return dict(seq_stck=stck, seq_stck_fact_eps=kv["STCK_FACT_EPS"], seq_hydr_AT=at, seq_hydr_GC=gc)


def run_reference(s: systems.System, T_str: str, salt: float, extra: dict | None = None, interaction="DNA2_nomesh",
                  step: int = 0, files: dict | None = None):
    """Step-0 observables of the reference binary on s.  step: the step
    counter written in the configuration (kept with restart_step_counter =
    false); files: extra input files (name -> text) written next to the
    input file, such as an external_forces_file."""
    extra = extra or {}
    s = s.copy()
    # zero velocities, so that the pressure observable holds the pair virial
    # only; a tiny angular momentum, which the reference requires without
    # refresh_vel (src/Backends/MDBackend.cpp:99-103) and which enters no
    # observable here
    s.vel = np.zeros((s.N, 3))
    s.L = np.zeros((s.N, 3))
    s.L[:, 0] = 1e-6
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        io.write_topology(os.path.join(d, "t.top"), s)
        io.write_conf(os.path.join(d, "c.dat"), s, step=step)
        for fname, text in (files or {}).items():
            with open(os.path.join(d, fname), "w") as f:
                f.write(text)
        if extra.get("use_average_seq") == "false":
            shutil.copy(SEQ_FILE, os.path.join(d, "seq.txt"))
            extra = dict(extra, seq_dep_file="seq.txt")
        inp = {
            "backend": "CPU", "sim_type": "MD", "interaction_type": interaction, "steps": "0", "seed": "1",
            "T": T_str, "salt_concentration": repr(salt), "dt": "0.003", "verlet_skin": "0.05",
            "thermostat": "no", "refresh_vel": "0", "topology": "t.top", "conf_file": "c.dat",
            "trajectory_file": "traj.dat", "energy_file": "energy.dat", "print_energy_every": "1000",
            "print_conf_interval": "100000", "no_stdout_energy": "1", "restart_step_counter": "0" if step else "1",
            "box_type": "orthogonal", "time_scale": "linear", "plugin_search_path": PLUGIN_DIR, "log_file": "log.dat",
        }
        inp.update({k: str(v) for k, v in extra.items()})
        with open(os.path.join(d, "input"), "w") as f:
            for k, v in inp.items():
                f.write(f"{k} = {v}\n")
            f.write("data_output_1 = {\n name = split.dat\n print_every = 1\n col_1 = {\n type = potential_energy\n"
                    " split = true\n precision = 15\n }\n}\n")
            f.write("data_output_2 = {\n name = forces.dat\n print_every = 1\n col_1 = {\n type = ForceDump\n }\n}\n")
            f.write("data_output_3 = {\n name = stress.dat\n print_every = 1\n col_1 = {\n type = pressure\n"
                    " stress_tensor = true\n custom_stress_tensor = true\n }\n}\n")
            f.write("data_output_4 = {\n name = hb.dat\n print_every = 1\n col_1 = {\n type = hb_list\n"
                    " only_count = true\n }\n}\n")
        r = subprocess.run([REF_BIN, "input"], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            log = open(os.path.join(d, "log.dat")).read() if os.path.exists(os.path.join(d, "log.dat")) else ""
            raise RuntimeError(f"reference failed: {r.stdout[-2000:]} {r.stderr[-2000:]} {log[-2000:]}")
        split = np.array([float(x) for x in open(os.path.join(d, "split.dat")).read().split()])
        # P, then the stress tensor xx yy zz xy xz yz divided by the volume
        stress = np.array([float(x) for x in open(os.path.join(d, "stress.dat")).read().split()])
        rows = np.loadtxt(os.path.join(d, "forces.dat"), ndmin=2)
        hb = int(open(os.path.join(d, "hb.dat")).read().split()[0])
    idx = rows[:, 0].astype(int)
    F = np.zeros((s.N, 3))
    T = np.zeros((s.N, 3))
    F[idx] = rows[:, 1:4]
    T[idx] = rows[:, 4:7]
    return split * s.N, F, T, stress, hb


def tangle(seed: int, n_strands: int = 6, length: int = 8, box: float = 6.0, bb: float = 0.70, ss: float = 0.36,
           bs: float = 0.53) -> systems.System:
    """Short strands grown as random walks of the backbone site with random
    nucleotide orientations, packed in a small box, so that HB, cross and
    coaxial stacking are sampled at non-ideal angles.  Candidates closer
    than bb (back-back), ss (base-base) or bs (back-base) to any other
    nucleotide are rejected; the bonded neighbour is exempt from back-back
    only.  The defaults keep base-base and back-base sites outside the
    excluded-volume cut-offs (EXCL_RC2 = 0.3354, EXCL_RC3 = EXCL_RC4 =
    0.5233, src/model.h:78-81) and back-back sites at the edge of the
    smoothing shell below EXCL_RC1 = 0.7119; thresholds below them let
    sites into the smoothing shells and the Lennard-Jones cores, bonded
    neighbours included."""
    rng = np.random.default_rng(seed)
    b = systems._Builder()
    placed = []          # (back, base) of every placed nucleotide

    def dist(p, q):
        d = p - q
        d -= box * np.rint(d / box)
        return np.sqrt(d @ d)

    def clear(back, base, skip_backback):
        for k, (qb, qs) in enumerate(placed):
            if k not in skip_backback and dist(back, qb) < bb:
                return False
            if dist(base, qs) < ss or dist(back, qs) < bs or dist(base, qb) < bs:
                return False
        return True

    for _ in range(n_strands):
        for attempt in range(2000):
            pos, a1s, a3s, mine = [], [], [], []
            back_prev = rng.random(3) * box
            good = True
            for k in range(length):
                for tries in range(400):
                    R = systems.random_rotation(rng)
                    a1, a3 = R[:, 0], R[:, 2]
                    a2 = np.cross(a3, a1)
                    if k == 0:
                        back = back_prev
                    else:
                        u = rng.normal(size=3)
                        u /= np.linalg.norm(u)
                        back = back_prev + u * rng.uniform(0.62, 0.9)
                    c = back - (systems.BACK1 * a1 + systems.BACK2 * a2)
                    base = c + 0.4 * a1
                    skip = {len(placed) + len(mine) - 1} if k > 0 else set()
                    saved = placed[:]
                    placed.extend(mine)
                    ok = clear(back, base, skip)
                    placed[:] = saved
                    if ok:
                        break
                else:
                    good = False
                    break
                mine.append((back, base))
                pos.append(c)
                a1s.append(a1)
                a3s.append(a3)
                back_prev = back
            if good:
                placed.extend(mine)
                b.add_strand(pos, a1s, a3s, rng.integers(0, 4, length))
                break
    return b.build([box] * 3, f"tangle{seed}")


def equilibrate(s: systems.System, T_str: str, salt: float, steps: int, seed: int = 7) -> systems.System:
    """Runs the reference binary for `steps` MD steps (Langevin, dt 0.003)
    from s and returns its last configuration: thermal configurations
    produced by the reference itself."""
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        io.write_topology(os.path.join(d, "t.top"), s)
        io.write_conf(os.path.join(d, "c.dat"), s)
        inp = {
            "backend": "CPU", "sim_type": "MD", "interaction_type": "DNA2_nomesh", "steps": str(steps),
            "seed": str(seed), "T": T_str, "salt_concentration": repr(salt), "dt": "0.003", "verlet_skin": "0.05",
            "thermostat": "langevin", "diff_coeff": "2.5", "refresh_vel": "1", "topology": "t.top",
            "conf_file": "c.dat", "trajectory_file": "traj.dat", "energy_file": "energy.dat",
            "print_energy_every": "1000", "print_conf_interval": "1000000", "no_stdout_energy": "1",
            "restart_step_counter": "1", "box_type": "orthogonal", "time_scale": "linear", "log_file": "log.dat",
            "lastconf_file": "last.dat", "fix_diffusion": "0",
        }
        with open(os.path.join(d, "input"), "w") as f:
            for k, v in inp.items():
                f.write(f"{k} = {v}\n")
        r = subprocess.run([REF_BIN, "input"], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("reference MD run failed: " + open(os.path.join(d, "log.dat")).read()[-2000:])
        out = io.load_system(os.path.join(d, "t.top"), os.path.join(d, "last.dat"), name=s.name + "_md")
    out.vel = out.L = None
    return out


def nicked_duplex(seed: int) -> systems.System:
    """A 24-bp duplex whose first strand is cut in the middle: the two
    nucleotides at the nick stack coaxially without being bonded."""
    rng = np.random.default_rng(seed)
    d = systems._Builder()
    systems.add_duplex(d, rng.integers(0, 4, 24), (10.0, 10.0, 10.0))
    s = d.build([20.0] * 3, "nicked")
    s.n5[11] = -1
    s.n3[12] = -1
    s.strand[12:24] = 2
    return s


def bundle_small(seed: int) -> systems.System:
    return systems.duplex_bundle(seed, nx=3, ny=2, bp=12, spacing=2.6, lz=12 * systems.BASE_BASE + 3.0)


def compressed_duplex(seed: int, bp: int = 12, scale: float = 0.78) -> systems.System:
    """An ideal duplex whose rise is scaled by `scale`: consecutive bases of
    a strand overlap inside the bonded excluded-volume ranges."""
    rng = np.random.default_rng(seed)
    b = systems._Builder()
    pos, a1, a3 = systems.duplex_geometry(bp, direction=(0.0, 0.0, 1.0), perp=(1.0, 0.0, 0.0))
    z = pos[:, 2].copy()
    pos[:, 2] = z * scale
    pos += np.array([5.0, 5.0, 5.0]) - pos.mean(axis=0)
    seq = list(rng.integers(0, 4, bp))
    b.add_strand(pos[:bp], a1[:bp], a3[:bp], seq)
    b.add_strand(pos[bp:], a1[bp:], a3[bp:], [3 - x for x in reversed(seq)])
    return b.build([10.0] * 3, "compressed")


def pbc_helix(seed: int) -> systems.System:
    """Four 12-bp duplexes on a hexagonal lattice in a box whose z side is
    exactly 12 rises: each helix continues through the periodic boundary,
    so its first and last base pairs stack coaxially without being
    bonded."""
    return systems.duplex_bundle(seed, nx=2, ny=2, bp=12, spacing=2.6, lz=12 * systems.BASE_BASE)


def custom_btype_duplex(seed: int) -> systems.System:
    """A 16-bp duplex whose first 10 pairs use custom base types
    (|btype| >= 300, complementary when the two sum to 3), the hb_multiplier
    path of src/Interactions/DNAInteraction.cpp:785-786."""
    rng = np.random.default_rng(seed)
    b = systems._Builder()
    systems.add_duplex(b, rng.integers(0, 4, 16), (8.0, 8.0, 8.0))
    s = b.build([16.0] * 3, "custom_btype")
    for k in range(10):
        x = int(s.btype[k])
        mate = 31 - k        # strand 2 is reversed: pair k sits at 2 bp - 1 - k
        assert s.btype[mate] == 3 - x
        s.btype[k] = 300 + x
        s.btype[mate] = 3 - (300 + x)
    return s


def md_frame(s: systems.System, T_str: str, salt: float, steps: int, every: int, col: int, seed: int = 7,
             skip: int = 0) -> systems.System:
    """Runs the reference binary (Langevin MD, dt 0.003) from s, reads every
    `every`-th configuration of its trajectory and returns a frame in which
    term `col` of the energy split is nonzero, preferring frames where more
    terms are; `skip` passes over that many qualifying frames first.  The
    FP64 oracle only selects the frame; the fixture's values come from the
    reference binary."""
    sys.path.insert(0, ROOT)
    from oracle import pyoracle as po
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        io.write_topology(os.path.join(d, "t.top"), s)
        io.write_conf(os.path.join(d, "c.dat"), s)
        inp = {
            "backend": "CPU", "sim_type": "MD", "interaction_type": "DNA2_nomesh", "steps": str(steps),
            "seed": str(seed), "T": T_str, "salt_concentration": repr(salt), "dt": "0.003", "verlet_skin": "0.05",
            "thermostat": "langevin", "diff_coeff": "2.5", "refresh_vel": "1", "topology": "t.top",
            "conf_file": "c.dat", "trajectory_file": "traj.dat", "energy_file": "energy.dat",
            "print_energy_every": "100000", "print_conf_interval": str(every), "no_stdout_energy": "1",
            "restart_step_counter": "1", "box_type": "orthogonal", "time_scale": "linear", "log_file": "log.dat",
            "fix_diffusion": "0",
        }
        with open(os.path.join(d, "input"), "w") as f:
            for k, v in inp.items():
                f.write(f"{k} = {v}\n")
        r = subprocess.run([REF_BIN, "input"], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("reference MD run failed: " + open(os.path.join(d, "log.dat")).read()[-2000:])
        lines = open(os.path.join(d, "traj.dat")).read().splitlines()
    P = po.Params(T=io.temperature(T_str), salt=salt)
    found = []
    for k in range(0, len(lines), s.N + 3):
        rows = np.array([[float(x) for x in l.split()] for l in lines[k + 3:k + 3 + s.N]])
        f = s.copy()
        f.pos, f.a1, f.a3 = rows[:, 0:3].copy(), rows[:, 3:6].copy(), rows[:, 6:9].copy()
        f.vel = f.L = None
        e = po.forces(P, f)["E8"]
        if abs(e[col]) > 1e-9:
            found.append((-(np.abs(e) > 1e-9).sum(), k, f))
    if len(found) <= skip:
        raise RuntimeError(f"no frame with term {col} nonzero")
    found.sort(key=lambda t: (t[0], t[1]))
    out = found[skip][2]
    out.name = s.name + "_md"
    return out


def hot_strands(seed: int) -> systems.System:
    """Three 16-nt single strands side by side in a 7.5 x 4 x 7 box."""
    rng = np.random.default_rng(seed)
    b = systems._Builder()
    for k in range(3):
        systems.add_single_strand(b, rng.integers(0, 4, 16), (1.25 + 2.5 * k, 2.0, 3.5))
    return b.build([7.5, 4.0, 7.0], "ss_hot")


def dense_duplexes(seed: int) -> systems.System:
    """Eight 8-bp duplexes in two layers of a 6 x 6 x 7.2 box."""
    rng = np.random.default_rng(seed)
    b = systems._Builder()
    for k in range(4):
        for l in range(2):
            systems.add_duplex(b, rng.integers(0, 4, 8), (1.5 + 3.0 * (k % 2), 1.5 + 3.0 * (k // 2), 1.7 + 3.6 * l))
    return b.build([6.0, 6.0, 7.2], "duplex_dense")


EXCL, BEXCL, CXST = 3, 1, 6
SEQDEP = {"use_average_seq": "false"}


def cases():
    """(name, builder, T, salt, extra input keys); builders run only for
    fixtures that do not exist yet."""
    def ff():
        s = io.load_system("/root/reference/test/DNA/FORCE_FIELD/init.top",
                           "/root/reference/test/DNA/FORCE_FIELD/init.dat")
        s.name = "ff16"
        return s

    def hp(k):
        return lambda: systems.hairpin_set(3, n=4)[k]

    yield "ff16_20C_1M", ff, "20C", 1.0, {}
    yield "ff16_30C_0.2M", ff, "30C", 0.2, {}
    yield "duplex20_20C_0.5M", lambda: systems.duplex_20bp(0), "20C", 0.5, {}
    yield "hairpin_a", hp(0), "25C", 0.5, {}
    yield "hairpin_b", hp(3), "75C", 0.5, {}
    yield "bundle_30C_0.2M", lambda: bundle_small(1), "30C", 0.2, {}
    yield "gas8_25C_0.5M", lambda: systems.duplex_gas(5, n_duplex=8, bp=24, density=0.04), "25C", 0.5, {}
    yield "duplex20_md_20C", lambda: equilibrate(systems.duplex_20bp(2), "20C", 0.5, 4000), "20C", 0.5, {}
    yield "nicked_md_30C", lambda: equilibrate(nicked_duplex(4), "30C", 0.5, 4000), "30C", 0.5, {}
    yield "hairpin_md_50C", lambda: equilibrate(hp(1)(), "50C", 0.5, 4000), "50C", 0.5, {}
    yield "bundle_md_30C", lambda: equilibrate(bundle_small(6), "30C", 0.2, 2000), "30C", 0.2, {}
    yield "tangle1_20C_0.5M", lambda: tangle(1), "20C", 0.5, {}
    yield "tangle2_45C_0.1M", lambda: tangle(2, n_strands=8, length=6, box=5.0), "45C", 0.1, {}
    yield "tangle3_nohalf_dh", lambda: tangle(3), "37C", 0.3, {"dh_half_charged_ends": "false", "dh_lambda": "0.4",
                                                                "dh_strength": "0.06"}
    # excluded volume (bonded and non-bonded) and coaxial stacking
    yield "compressed_20C", lambda: compressed_duplex(1), "20C", 0.5, {}
    yield "tangle4_dense_30C", lambda: tangle(4, bb=0.62, ss=0.29, bs=0.46), "30C", 0.5, {}
    yield "tangle6_dense_60C_0.2M", lambda: tangle(6, bb=0.62, ss=0.29, bs=0.46), "60C", 0.2, {}
    yield "pbc_helix_30C_0.2M", lambda: pbc_helix(1), "30C", 0.2, {}
    yield "hot_md_bexcl_95C", lambda: md_frame(hot_strands(11), "95C", 0.5, 40000, 1000, BEXCL), "95C", 0.5, {}
    yield "hot_md_cxst_95C", lambda: md_frame(hot_strands(11), "95C", 0.5, 40000, 1000, CXST), "95C", 0.5, {}
    yield "dense_md_excl_30C", lambda: md_frame(dense_duplexes(12), "30C", 0.5, 40000, 1000, EXCL), "30C", 0.5, {}
    # interaction options
    yield "mbf_md_20C", lambda: equilibrate(systems.duplex_20bp(5), "20C", 0.5, 4000), "20C", 0.5, \
        {"max_backbone_force": "0.3", "max_backbone_force_far": "0.05"}
    yield "mbf_hairpin_60C", lambda: equilibrate(hp(2)(), "60C", 0.5, 4000), "60C", 0.5, {"max_backbone_force": "1.0"}
    yield "hbmult_custom_25C", lambda: equilibrate(custom_btype_duplex(3), "25C", 0.5, 2000), "25C", 0.5, \
        {"hb_multiplier": "1.5"}
    yield "dh_rhigh_bundle_30C", lambda: bundle_small(2), "30C", 0.2, {"debye_huckel_rhigh": "1.1"}
    yield "seqdep_md_25C", lambda: equilibrate(systems.duplex_20bp(6), "25C", 0.5, 4000), "25C", 0.5, SEQDEP
    yield "seqdep_hairpin_45C", lambda: equilibrate(hp(0)(), "45C", 0.5, 3000), "45C", 0.5, SEQDEP


def option_fields(extra: dict) -> dict:
    """The input-file options of a case as fixture fields (defaults where
    absent: -1 means the reference's own default)."""
    out = dict(half_charged_ends=int(extra.get("dh_half_charged_ends", "true") != "false"),
               dh_lambda=float(extra.get("dh_lambda", -1.0)), dh_strength=float(extra.get("dh_strength", -1.0)),
               dh_rhigh=float(extra.get("debye_huckel_rhigh", -1.0)),
               hb_multiplier=float(extra.get("hb_multiplier", 1.0)),
               max_backbone_force=float(extra.get("max_backbone_force", -1.0)),
               max_backbone_force_far=float(extra.get("max_backbone_force_far", -1.0)),
               seqdep=int(extra.get("use_average_seq") == "false"))
    if out["seqdep"]:
        out.update(read_seq_file(SEQ_FILE))
    return out


def stored_system(z, name) -> systems.System:
    return systems.System(pos=z["pos"].copy(), a1=z["a1"].copy(), a3=z["a3"].copy(), btype=z["btype"].copy(),
                          n3=z["n3"].copy(), n5=z["n5"].copy(), strand=z["strand"].copy(), box=z["box"].copy(),
                          name=name)


def main():
    if not os.path.exists(REF_BIN):
        raise SystemExit("build the reference first: make -f oracle/ref.mk -j8")
    only = set(sys.argv[1:])
    for name, build, T_str, salt, extra in cases():
        if only and name not in only:
            continue
        path = os.path.join(HERE, f"{name}.npz")
        old = dict(np.load(path)) if os.path.exists(path) else None
        s = stored_system(old, name) if old is not None else build()
        E, F, T, stress, hb = run_reference(s, T_str, salt, extra)
        a2 = np.cross(s.a3, s.a1)
        rec = dict(pos=s.pos, a1=s.a1, a2=a2, a3=s.a3, btype=s.btype, n3=s.n3, n5=s.n5, strand=s.strand, box=s.box,
                   T=io.temperature(T_str), salt=salt, E8=E, F=F, Tbody=T, stress7=stress, hb_count=hb)
        rec.update(option_fields(extra))
        if old is not None:
            for k in ("E8", "F", "Tbody", "stress7"):
                if not np.array_equal(old[k], rec[k]):
                    raise RuntimeError(f"{name}: the reference binary no longer reproduces the stored {k}")
            old.update({k: v for k, v in rec.items() if k not in old})
            rec = old
        np.savez_compressed(path, **rec)
        nz = "".join("x" if abs(e) > 1e-9 else "." for e in E)
        print(f"{name}: N={s.N} U/N={E.sum() / s.N:.6f} max|F|={np.abs(F).max():.3g} terms {nz} hb {hb}")


if __name__ == "__main__":
    main()
