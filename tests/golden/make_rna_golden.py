"""Writes the oxRNA fixtures of tests/golden/rna/ with the reference binary
(run in the build container, after `make -f oracle/ref.mk`):

    python tests/golden/make_rna_golden.py [case names]

Each case is a state, an interaction (RNA2, or RNA without Debye-Hueckel),
a temperature, a salt concentration and options.  The reference CPU
backend (steps = 0) writes the energy split (potential_energy, split =
true), the forces and torques (ForceDump observable plugin, %.17g), the
stress tensor (pressure observable) and the hydrogen-bond count (hb_list,
only_count), as tests/golden/make_golden.py does for oxDNA2; the .npz holds
inputs and outputs only.

The reference's CPU backend interpolates the hydrogen-bonding f4 factors on
cubic meshes (src/Interactions/RNAInteraction.cpp:312-324, 769-793); the
GPU kernel evaluates them analytically, as the reference's own GPU kernel
does (src/CUDA/Interactions/CUDA_RNA.cuh:743-761).  So every case with
hydrogen bonds has a twin "_nohb" with every base type replaced by one that
pairs with none (all A, btype 0: A-A sums to 0, not 3): the same state with
the hydrogen-bonding term at zero, where every other term is the
reference's to FP64 rounding.

States: the reference's own oxRNA force-field test (test/RNA/FORCE_FIELD:
16 nt in three strands, 20 C, 1 M; its published energy split
AVG_SEQ/reference.dat is kept as E8_published), the relaxed 8-bp duplex of
examples/RNA_DUPLEX_MELT (custom base types 13..20 / -17..-10, unwrapped
coordinates far from the box), generated A-form duplexes
(oxdna_amd.systems.add_rna_duplex: a 40-bp one, a nicked one for coaxial
stacking), and thermal states that the reference binary's own Langevin MD
makes from these (hot duplexes, a dense box of 8-bp duplexes and a frame of
it with non-bonded contacts and coaxial stacks, collapsed single strands
with bonded excluded volume).
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
sys.path.insert(0, HERE)

import make_golden as mg  # noqa: E402
from oxdna_amd import io, systems  # noqa: E402

OUT = os.path.join(HERE, "rna")
REF = "/root/reference"
FF_DIR = os.path.join(REF, "test/RNA/FORCE_FIELD")
MELT_DIR = os.path.join(REF, "examples/RNA_DUPLEX_MELT")


def ref_md(s: systems.System, interaction: str, T_str: str, salt: float, steps: int, seed: int = 11) -> systems.System:
    """The last configuration of `steps` steps of the reference binary's
    Langevin MD (dt 0.003, diff_coeff 2.5) from s."""
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        io.write_topology(os.path.join(d, "t.top"), s)
        io.write_conf(os.path.join(d, "c.dat"), s)
        inp = {
            "backend": "CPU", "sim_type": "MD", "interaction_type": interaction, "steps": str(steps),
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
        r = subprocess.run([mg.REF_BIN, "input"], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("reference MD run failed: " + open(os.path.join(d, "log.dat")).read()[-2000:])
        out = io.load_system(os.path.join(d, "t.top"), os.path.join(d, "last.dat"), name=s.name + "_md")
    out.vel = out.L = None
    return out


def ref_md_pick(s: systems.System, T_str: str, salt: float, steps: int, every: int, cols, seed: int = 11) -> systems.System:
    """Runs the reference binary's Langevin MD (RNA2) from s, evaluates
    every `every`-th configuration of its trajectory with the reference
    binary and returns the first whose energy split is nonzero in every
    column of `cols`."""
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        io.write_topology(os.path.join(d, "t.top"), s)
        io.write_conf(os.path.join(d, "c.dat"), s)
        inp = {
            "backend": "CPU", "sim_type": "MD", "interaction_type": "RNA2", "steps": str(steps),
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
        r = subprocess.run([mg.REF_BIN, "input"], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("reference MD run failed: " + open(os.path.join(d, "log.dat")).read()[-2000:])
        lines = open(os.path.join(d, "traj.dat")).read().splitlines()
    for k in range(0, len(lines), s.N + 3):
        rows = np.array([[float(x) for x in l.split()] for l in lines[k + 3:k + 3 + s.N]])
        f = s.copy()
        f.pos, f.a1, f.a3 = rows[:, 0:3].copy(), rows[:, 3:6].copy(), rows[:, 6:9].copy()
        f.vel = f.L = None
        E = mg.run_reference(f, T_str, salt, interaction="RNA2")[0]
        if all(abs(E[c]) > 1e-9 for c in cols):
            f.name = s.name + "_md"
            return f
    raise RuntimeError(f"no frame with terms {cols} nonzero")


def ff16() -> systems.System:
    s = io.load_system(os.path.join(FF_DIR, "init.top"), os.path.join(FF_DIR, "init.dat"), name="rna_ff16")
    s.vel = s.L = None
    return s


def melt8() -> systems.System:
    s = io.load_system(os.path.join(MELT_DIR, "sim.top"), os.path.join(MELT_DIR, "init.conf"), name="rna_melt8")
    s.vel = s.L = None
    return s


def duplex40() -> systems.System:
    rng = np.random.default_rng(3)
    b = systems._Builder()
    systems.add_rna_duplex(b, rng.integers(0, 4, 40), (15.0, 15.0, 15.0), systems.random_rotation(rng))
    return b.build([30.0, 30.0, 30.0], "rna_duplex40", T=systems.celsius(37.0), salt=0.5)


def dense8() -> systems.System:
    """27 random-sequence 8-bp duplexes with random orientations on a
    3 x 3 x 3 lattice of spacing 4.2 (box 12.6), run 4000 steps at 45 C."""
    rng = np.random.default_rng(5)
    b = systems._Builder()
    for ix in range(3):
        for iy in range(3):
            for iz in range(3):
                c = (np.array([ix, iy, iz]) + 0.5) * 4.2
                systems.add_rna_duplex(b, rng.integers(0, 4, 8), c, systems.random_rotation(rng))
    s = b.build([12.6, 12.6, 12.6], "rna_dense8", T=systems.celsius(45.0), salt=1.0)
    return ref_md(s, "RNA2", "45C", 1.0, 4000)


def dense8_contacts() -> systems.System:
    """The dense box of 8-bp duplexes after more steps at 45 C: the first
    frame (every 500 steps) with non-bonded excluded volume and coaxial
    stacking both acting."""
    return ref_md_pick(dense8(), "45C", 1.0, 40000, 500, (3, 6))


def hot_melt() -> systems.System:
    """Four generated 12-bp duplexes in a box of 14, run 20000 steps at
    95 C: partly melted strands."""
    rng = np.random.default_rng(9)
    b = systems._Builder()
    for k in range(4):
        c = np.array([3.5 + 7.0 * (k % 2), 3.5 + 7.0 * (k // 2), 7.0])
        systems.add_rna_duplex(b, rng.integers(0, 4, 12), c, systems.random_rotation(rng))
    s = b.build([14.0, 14.0, 14.0], "rna_hot", T=systems.celsius(95.0), salt=1.0)
    return ref_md(s, "RNA2", "95C", 1.0, 20000)


def nicked24() -> systems.System:
    """A 24-bp duplex whose strand 1 is cut between its 12th and 13th
    nucleotides: the two nucleotides at the nick stack coaxially, as
    non-bonded neighbours."""
    rng = np.random.default_rng(13)
    b = systems._Builder()
    systems.add_rna_duplex(b, rng.integers(0, 4, 24), (10.0, 10.0, 10.0), systems.random_rotation(rng))
    s = b.build([20.0, 20.0, 20.0], "rna_nicked24", T=systems.celsius(37.0), salt=1.0)
    # strand 1 holds rows 0..23, 3' end first: cut the bond 11 - 12
    s.n3[12] = -1
    s.n5[11] = -1
    s.strand = s.strand.copy()
    s.strand[12:24] = 2
    return s


def tangle12() -> systems.System:
    """Eight 12-nt single strands (one strand of a generated duplex each,
    random sequences and orientations) in a box of 8, run 10000 steps at
    60 C: collapsed strands with contacts, coaxial stacks and some base
    pairs."""
    rng = np.random.default_rng(19)
    b = systems._Builder()
    L = 8.0
    placed = np.zeros((0, 3))
    for k in range(8):
        pos, a1, a3 = systems.rna_duplex_geometry(12)
        c = (np.array([k % 2, (k // 2) % 2, k // 4]) + 0.5) * 4.0
        # orientations drawn until the strand keeps 1.2 from the others
        while True:
            R = systems.random_rotation(rng)
            p = pos[:12] @ R.T + c
            d = p[:, None, :] - placed[None, :, :]
            d -= L * np.rint(d / L)
            if placed.shape[0] == 0 or np.linalg.norm(d, axis=2).min() > 1.2:
                break
        placed = np.vstack([placed, p])
        seq = rng.integers(0, 4, 12)
        b.add_strand(p, a1[:12] @ R.T, a3[:12] @ R.T, list(seq))
    s = b.build([L, L, L], "rna_tangle12", T=systems.celsius(60.0), salt=1.0)
    return ref_md(s, "RNA2", "60C", 1.0, 10000)


def nohb(s: systems.System, name: str) -> systems.System:
    o = s.copy()
    o.btype = np.zeros_like(s.btype)
    o.name = name
    return o


def published_ff16() -> np.ndarray:
    """test/RNA/FORCE_FIELD/AVG_SEQ/reference.dat: the reference's energy
    split per nucleotide, 6 decimals."""
    return np.array([float(x) for x in open(os.path.join(FF_DIR, "AVG_SEQ/reference.dat")).read().split()])


# name -> (builder, interaction, T, salt, extra input keys)
def cases():
    return [
        ("rna_ff16", ff16, "RNA2", "20C", 1.0, {}),
        ("rna_ff16_nohb", lambda: nohb(ff16(), "rna_ff16_nohb"), "RNA2", "20C", 1.0, {}),
        ("rna_ff16_rna1", ff16, "RNA", "20C", 1.0, {}),
        ("rna_melt8_62C", melt8, "RNA2", "62C", 1.0, {}),
        ("rna_melt8_nohb_0.1M", lambda: nohb(melt8(), "rna_melt8_nohb"), "RNA2", "62C", 0.1, {"dh_half_charged_ends": "false"}),
        ("rna_duplex40_37C_0.5M", duplex40, "RNA2", "37C", 0.5, {}),
        ("rna_duplex40_nohb", lambda: nohb(duplex40(), "rna_duplex40_nohb"), "RNA2", "37C", 0.5, {}),
        ("rna_duplex40_opts", duplex40, "RNA2", "25C", 0.3, {"dh_lambda": "0.35", "dh_strength": "0.1", "debye_huckel_rhigh": "1.0"}),
        ("rna_dense8_45C", dense8, "RNA2", "45C", 1.0, {}),
        ("rna_dense8_nohb", lambda: nohb(dense8(), "rna_dense8_nohb"), "RNA2", "45C", 1.0, {}),
        ("rna_nicked24_37C", nicked24, "RNA2", "37C", 1.0, {}),
        ("rna_nicked24_nohb", lambda: nohb(nicked24(), "rna_nicked24_nohb"), "RNA2", "37C", 1.0, {}),
        ("rna_contacts_45C", dense8_contacts, "RNA2", "45C", 1.0, {}),
        ("rna_contacts_nohb", lambda: nohb(dense8_contacts(), "rna_contacts_nohb"), "RNA2", "45C", 1.0, {}),
        ("rna_tangle12_60C", tangle12, "RNA2", "60C", 1.0, {}),
        ("rna_tangle12_nohb", lambda: nohb(tangle12(), "rna_tangle12_nohb"), "RNA2", "60C", 1.0, {}),
        ("rna_hot_95C", hot_melt, "RNA2", "95C", 1.0, {}),
        ("rna_hot_nohb_mbf", lambda: nohb(hot_melt(), "rna_hot_nohb"), "RNA2", "95C", 1.0,
         {"max_backbone_force": "0.3", "max_backbone_force_far": "0.05"}),
    ]


def main(only=None):
    if not os.path.exists(mg.REF_BIN):
        raise SystemExit("build the reference first: make -f oracle/ref.mk -j8")
    os.makedirs(OUT, exist_ok=True)
    for name, build, inter, T_str, salt, extra in cases():
        if only and name not in only:
            continue
        path = os.path.join(OUT, f"{name}.npz")
        old = dict(np.load(path)) if os.path.exists(path) else None
        s = mg.stored_system(old, name) if old is not None else build()
        E, F, T, stress, hb = mg.run_reference(s, T_str, salt, extra, interaction=inter)
        rec = dict(pos=s.pos, a1=s.a1, a3=s.a3, btype=s.btype, n3=s.n3, n5=s.n5, strand=s.strand, box=s.box,
                   T=io.temperature(T_str), salt=salt, interaction=inter, E8=E, F=F, Tbody=T, stress7=stress,
                   hb_count=hb)
        rec.update(mg.option_fields(extra))
        if name == "rna_ff16":
            rec["E8_published"] = published_ff16()
        if old is not None:
            for k in ("E8", "F", "Tbody", "stress7"):
                if not np.array_equal(old[k], rec[k]):
                    raise RuntimeError(f"{name}: the reference binary no longer reproduces the stored {k}")
        np.savez_compressed(path, **rec)
        nz = "".join("x" if abs(e) > 1e-9 else "." for e in E)
        print(f"{name}: N={s.N} U/N={E.sum() / s.N:.6f} max|F|={np.abs(F).max():.3g} terms {nz} hb {hb}")


if __name__ == "__main__":
    main(set(sys.argv[1:]))
