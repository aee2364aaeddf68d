"""Host driver: runs an oxDNA input file (sim_type = MD, interaction DNA2,
RNA2 or RNA) on the HIP batch, writing the reference's output files.

    python -m oxdna_amd.driver input [key=value ...]

The run loop follows SimManager::run (src/Managers/SimManager.cpp:142-201):
at every step the observables due at that step are printed before the step
is taken (the energy file every print_energy_every steps, step 0 included),
the trajectory gets a configuration every print_conf_interval steps (not at
step 0), strands are re-boxed every fix_diffusion_every steps, and after the
last step the observables due then and the last configuration are printed.
Between output points the batch runs uninterrupted on the device (HIP graph
replay), so a run with sparse output never returns to the host.

Output formats are the reference's:
  - energy file: the step observable in MD units (`%14.4lf` of step * dt,
    src/Observables/Step.cpp:20-23), then potential, kinetic and total
    energy per particle (`%10.6lf`, src/Observables/TotalEnergy.cpp:32-38,
    src/Observables/BaseObservable.cpp:32-43), space separated;
  - trajectory and last configuration: `t = `, `b = `, `E = ` headers and
    one line of position, a1, a3, velocity, angular momentum per nucleotide
    (oxdna_amd/io.py).

Input keys read (reference key names): topology, conf_file, steps, T, dt,
salt_concentration, verlet_skin, thermostat, newtonian_steps, diff_coeff,
gamma_trans, pt, bussi_tau, refresh_vel, seed, restart_step_counter,
backend_precision, dh_lambda, dh_strength, debye_huckel_rhigh,
dh_half_charged_ends, hb_multiplier, max_backbone_force,
max_backbone_force_far, energy_file, print_energy_every, trajectory_file,
print_conf_interval, lastconf_file, lastconf_file_bin, binary_initial_conf,
fix_diffusion, fix_diffusion_every, CUDA_sort_every, CUDA_device,
checkpoint_every, checkpoint_file, reload_from, back_in_box,
use_average_seq with seq_dep_file, external_forces with
external_forces_file and external_forces_as_JSON, data_output_N (class
DataOutput: step, potential_energy, kinetic_energy and hb_list
observables).  Binary configurations are
the reference's (oxdna_amd/io.py: write_conf_binary, read_conf_binary).
Configurations are written with absolute (unwrapped) positions, the stored
position plus its periodic image count times the box, unless back_in_box
(src/Observables/Configurations/Configuration.cpp:110-135); binary ones
carry the stored position and the image counts.  Without a seed (or with
seed = 0) the seed comes from the clock and /dev/urandom, as in the
reference (src/Managers/SimManager.cpp:78-90), and is printed.
A checkpoint file holds the step (one JSON line) and the batch's whole
state (oxh_batch_save); reload_from continues from it bitwise, the batch
being built from topology and conf_file first (here conf_file only shapes
the batch).  reload_from also takes the reference's own reload file, a
binary configuration, with the reference's rules: no conf_file, no seed,
restart_step_counter false.  sim_type must be MD and interaction_type
DNA2 or DNA2_nomesh (the GPU interaction is the analytic DNA2_nomesh form),
RNA2 or RNA (oxRNA with the analytic hydrogen-bonding f4 of the reference's
GPU kernel; average sequence, no mismatch_repulsion).
"""
from __future__ import annotations

import os
import sys

import numpy as np

from . import io

_BOOL_TRUE = {"1", "true", "yes", "on"}


def _bool(v) -> bool:
    return str(v).strip().lower() in _BOOL_TRUE


def _num(inp: dict, key: str, cast=float, default=None):
    if key not in inp:
        if default is None:
            raise KeyError(f"input key '{key}' is required")
        return default
    return cast(float(inp[key])) if cast is int else cast(inp[key])


def random_seed() -> int:
    """The reference's seed when none is given: the clock plus an int from
    /dev/urandom (src/Managers/SimManager.cpp:78-90)."""
    import time
    seed = int(time.time())
    try:
        seed += int.from_bytes(os.urandom(4), "little", signed=True)
    except (NotImplementedError, OSError):
        pass
    return seed & 0xFFFFFFFF


def seq_dep_options(path: str) -> dict:
    """Sequence-dependent well depths of a seq_dep_file, read into FP32 as
    DNAInteraction::init reads them (getInputFloat,
    src/Interactions/DNAInteraction.cpp:328-376): STCK_X_Y for the 16 base
    pairs, STCK_FACT_EPS, HYDR_A_T (or HYDR_T_A) and HYDR_G_C (or
    HYDR_C_G)."""
    import numpy as np
    with open(path) as f:
        kv = {k: float(np.float32(float(v))) for k, v in io.parse_input(f.read()).items()}
    names = "AGCT"
    try:
        stck = [[kv[f"STCK_{names[i]}_{names[j]}"] for j in range(4)] for i in range(4)]
        at = kv["HYDR_A_T"] if "HYDR_A_T" in kv else kv["HYDR_T_A"]
        gc = kv["HYDR_G_C"] if "HYDR_G_C" in kv else kv["HYDR_C_G"]
        fact = kv["STCK_FACT_EPS"]
    except KeyError as e:
        raise ValueError(f"seq_dep_file {path}: missing key {e}") from None
    return dict(seq_stck=stck, seq_stck_fact_eps=fact, seq_hydr_AT=at, seq_hydr_GC=gc)


def batch_options(inp: dict, path=lambda p: p) -> dict:
    """HIPBatch options (reference key names) from an input dict; `path`
    resolves the file names the input gives (seq_dep_file)."""
    st = inp.get("sim_type", "MD")
    if st != "MD":
        raise ValueError(f"sim_type = {st}: this driver runs molecular dynamics only")
    it = inp.get("interaction_type", "DNA")
    if it not in ("DNA2", "DNA2_nomesh", "RNA2", "RNA"):
        raise ValueError(f"interaction_type = {it}: this driver runs the oxDNA2 (DNA2) and oxRNA (RNA, RNA2) interactions")
    rna = it in ("RNA2", "RNA")
    if rna and _bool(inp.get("mismatch_repulsion", "false")):
        raise ValueError("mismatch_repulsion = true is not available")
    if rna and "seq_dep_file_RNA" in inp and not _bool(inp.get("use_average_seq", "true")):
        raise ValueError("the oxRNA models run with the average-sequence parameters (use_average_seq = true)")
    th = inp.get("thermostat", "no")
    if th == "john":
        th = "brownian"
    opts = dict(
        interaction_type=it,
        T=io.temperature(inp["T"]),
        dt=_num(inp, "dt"),
        # required by DNA2; RNA2 defaults to 1 M (src/Interactions/RNAInteraction2.cpp:36-38)
        salt_concentration=_num(inp, "salt_concentration", float, 1.0) if rna else _num(inp, "salt_concentration"),
        verlet_skin=_num(inp, "verlet_skin", float, 0.05),
        thermostat=th,
        refresh_vel=_bool(inp.get("refresh_vel", "false")),
        seed=_num(inp, "seed", int, 0),
        backend_precision=inp.get("backend_precision", "mixed"),
        dh_half_charged_ends=_bool(inp.get("dh_half_charged_ends", "true")),
    )
    for key, cast in (("newtonian_steps", int), ("bussi_tau", int), ("diff_coeff", float), ("gamma_trans", float),
                      ("pt", float), ("dh_lambda", float), ("dh_strength", float), ("debye_huckel_rhigh", float),
                      ("hb_multiplier", float), ("max_backbone_force", float), ("max_backbone_force_far", float)):
        if key in inp:
            opts[key] = _num(inp, key, cast)
    if not _bool(inp.get("use_average_seq", "true")):
        if rna:
            raise ValueError("the oxRNA models run with the average-sequence parameters (use_average_seq = true)")
        if "seq_dep_file" not in inp:
            raise ValueError("use_average_seq = false needs seq_dep_file")
        opts.update(seq_dep_options(path(inp["seq_dep_file"])))
    if opts["seed"] == 0:
        opts["seed"] = random_seed()
    return opts


def output_steps(start: int, steps: int, energy_every: int, conf_every: int) -> list:
    """(step, print_energy, print_conf) for the steps of a run of `steps`
    steps from step `start` at which something is printed, in order, as
    SimManager::run does: energy at every multiple of energy_every from
    start to start + steps inclusive; configurations at multiples of
    conf_every after start, up to start + steps inclusive."""
    out = {}
    end = start + steps
    if energy_every > 0:
        first = start + (-start) % energy_every
        for t in range(first, end + 1, energy_every):
            out.setdefault(t, [False, False])[0] = True
    if conf_every > 0:
        first = start + (-start) % conf_every
        if first == start:
            first += conf_every
        for t in range(first, end + 1, conf_every):
            out.setdefault(t, [False, False])[1] = True
    return [(t, e, c) for t, (e, c) in sorted(out.items())]


def energy_line(step: int, dt: float, U: float, K: float) -> str:
    return f"{step * dt:14.4f} {U:10.6f} {K:10.6f} {U + K:10.6f}"


def _block(text: str) -> dict:
    """key = value pairs inside one { ... } block of an input file."""
    body = text.strip()
    if not (body.startswith("{") and body.endswith("}")):
        raise ValueError(f"malformed block: {text[:60]!r}")
    return io.parse_input(body[1:-1])


class DataOutput:
    """One data_output_N block (src/Observables/ObservableOutput.cpp:23-80,
    205-241): columns printed on one line, separated by a space, at every
    multiple of print_every in [start_from, stop_at].  Observables:
    step (src/Observables/Step.cpp: "%12lld", or "%14.4lf" of step * dt
    with units = MD), potential_energy (per nucleotide, or the 8-term split
    with split = true, each term preceded by a space;
    src/Observables/PotentialEnergy.cpp:21-55), kinetic_energy
    (src/Observables/KineticEnergy.cpp:37-50) and hb_list with only_count
    (pairs with HB energy below -0.1; src/Observables/HBList.cpp:141-164)
    and order_parameters (src/Observables/OrderParameterValues.cpp:22-80:
    each bond parameter's count of pairs with HB energy below -0.1, then
    each mindistance parameter's distance, every value followed by a space,
    in the stream's default format; the file from order_parameters_file or
    op_file in the column block, else in the input file).
    Numbers print with "%10.<precision>lf", or "%.<precision>g" with
    general_format (src/Observables/BaseObservable.cpp:32-43)."""

    TYPES = ("step", "potential_energy", "kinetic_energy", "hb_list", "order_parameters")

    def __init__(self, text: str, dt: float, path, inp: dict | None = None, N: int = 0):
        d = _block(text)
        if "name" not in d:
            raise ValueError("data_output block without a name")
        self.name = path(d["name"])
        self.every = int(float(d.get("print_every", "0")))
        self.start_from = int(float(d.get("start_from", "0")))
        self.stop_at = int(float(d.get("stop_at", "-1")))
        if not _bool(d.get("linear", "true")):
            raise ValueError("data_output with linear = false (log-spaced output) is not supported")
        self.dt = dt
        self.cols = []
        k = 1
        while f"col_{k}" in d:
            c = _block(d[f"col_{k}"])
            t = c.get("type", "")
            if t not in self.TYPES:
                raise ValueError(f"observable type {t!r} is not supported (supported: {', '.join(self.TYPES)})")
            if t == "hb_list" and not _bool(c.get("only_count", "false")):
                raise ValueError("hb_list is supported with only_count = true")
            prec = int(float(c.get("precision", "6")))
            fmt = f"{{:.{prec}g}}" if _bool(c.get("general_format", "false")) else f"{{:10.{prec}f}}"
            col = dict(type=t, fmt=fmt, split=_bool(c.get("split", "false")), md_units=c.get("units", "steps") == "MD")
            if t == "order_parameters":
                src = next((x for x in (c, inp or {}) for k in ("order_parameters_file", "op_file") if k in x), None)
                if src is None:
                    raise KeyError("order_parameters needs order_parameters_file or op_file")
                fname = src.get("order_parameters_file", src.get("op_file"))
                col["params"] = io.read_order_parameters(path(fname), N)
            self.cols.append(col)
            k += 1
        self.f = open(self.name, "w")

    def steps(self, start: int, end: int) -> list:
        if self.every < 1:
            return []
        first = max(start, self.start_from)
        first += (-first) % self.every
        last = end if self.stop_at < 0 else min(end, self.stop_at)
        return list(range(first, last + 1, self.every))

    def line(self, step: int, b, N: int) -> str:
        out = []
        for c in self.cols:
            t = c["type"]
            if t == "step":
                out.append(f"{step * self.dt:14.4f}" if c["md_units"] else f"{step:12d}")
            elif t == "potential_energy":
                if c["split"]:
                    out.append("".join(" " + c["fmt"].format(e / N) for e in b.energy_split(0)))
                else:
                    out.append(c["fmt"].format(b.energies()[0, 0] / N))
            elif t == "kinetic_energy":
                E = b.energies()[0]
                out.append(c["fmt"].format((E[2] + E[3]) / N))
            elif t == "hb_list":
                out.append(str(int(b.hb_count()[0])))
            else:
                hb, dist = order_parameter_values(b, c["params"])
                out.append("".join(f"{v} " for v in hb) + "".join(f"{v:g} " for v in dist))
        return " ".join(out)

    def close(self):
        self.f.close()


def order_parameter_values(b, params: list, sys: int = 0):
    """Values of the order parameters `params` (io.read_order_parameters)
    at the batch's current state, from the GPU's per-pair quantities: the
    count of each bond parameter's pairs with hydrogen-bonding energy below
    -0.1 (OrderParameterValues::get_output_string,
    src/Observables/OrderParameterValues.cpp:55-66) and each mindistance
    parameter's smallest base-base (or, with use_COM, centre-centre)
    distance (MinDistanceParameter::calculate_state,
    src/Utilities/OrderParameters.cpp:73-98).  Returns (counts, distances)
    in the file's order of each kind."""
    hb, dist = [], []
    for p in params:
        pairs = np.array(p["pairs"], dtype=np.int32).reshape(-1, 2)
        if p["type"] == "bond":
            e = b.pair_quantities("hb_energy", pairs, sys) if len(pairs) else np.zeros(0)
            hb.append(int(np.count_nonzero(e < io.HB_CUTOFF)))
        else:
            d = b.pair_quantities("centre_distance" if p["use_COM"] else "base_distance", pairs, sys) if len(pairs) else None
            dist.append(float(d.min()) if d is not None else -1.0)
    return hb, dist


def barostat_plan(inp: dict, seed: int, start: int, end: int) -> dict:
    """The MC barostat of an MD input (use_barostat, P, delta_L,
    barostat_probability, barostat_isotropic, barostat_molecular;
    src/Backends/MDBackend.cpp:43-60): the reference tries a volume move in
    each step with probability barostat_probability
    (MDBackend::_is_barostat_active, :121-126); here the steps of the tries
    in (start, end] are drawn up front as geometric gaps, and each try takes
    four uniform numbers, from a generator seeded by the run's seed."""
    out = dict(steps=set(), attempts=0, accepted=0)
    if not _bool(inp.get("use_barostat", "false")):
        return out
    for key in ("P", "delta_L", "barostat_probability"):
        if key not in inp:
            raise KeyError(f"use_barostat needs {key}")
    p = float(inp["barostat_probability"])
    if not 0.0 < p <= 1.0:
        raise ValueError("barostat_probability must be in (0, 1]")
    rng = np.random.default_rng([int(seed) & 0xFFFFFFFFFFFFFFFF, 0xBA80])
    t = start
    while True:
        t += int(rng.geometric(p))
        if t > end:
            break
        out["steps"].add(t)
    out.update(P=float(inp["P"]), delta_L=float(inp["delta_L"]), rng=rng,
               isotropic=_bool(inp.get("barostat_isotropic", "true")), molecular=_bool(inp.get("barostat_molecular", "false")))
    return out


def run(input_path: str, overrides=(), workdir: str | None = None) -> dict:
    """Runs the input file; returns the energy lines and the final state.
    Paths in the input are relative to workdir (default: the input file's
    directory)."""
    from .backend import HIPBatch

    import json

    inp = io.read_input(input_path, overrides)
    wd = workdir or os.path.dirname(os.path.abspath(input_path))
    path = lambda p: p if os.path.isabs(p) else os.path.join(wd, p)  # noqa: E731
    opts = batch_options(inp, path)
    if "seed" not in inp or _num(inp, "seed", int) == 0:
        print(f"seeding the RNG with {opts['seed']}", file=sys.stderr)
    restart = _bool(inp.get("restart_step_counter", "false"))
    # reload_from: this library's checkpoint (a JSON line, then the batch's
    # state), or the reference's own form, a binary configuration
    # (src/Backends/SimBackend.cpp:127-150): conf_file and seed are then not
    # given and the step counter is kept; the configuration holds positions,
    # orientations and momenta, and its stored generator words seed Philox
    reload_path = path(inp["reload_from"]) if "reload_from" in inp else None
    ckpt_reload = False
    if reload_path:
        with open(reload_path, "rb") as f:
            ckpt_reload = f.read(1) == b"{"
    if reload_path and not ckpt_reload:
        for key in ("conf_file", "seed"):
            if key in inp:
                raise ValueError(f"{key} cannot be specified together with reload_from (a binary configuration)")
        if restart:
            raise ValueError("restart_step_counter must be false when reload_from is specified")
        conf_path, binary = reload_path, True
        opts["refresh_vel"] = False
    else:
        conf_path, binary = path(inp["conf_file"]), _bool(inp.get("binary_initial_conf", "false"))
    s = io.load_system(path(inp["topology"]), conf_path, T=opts["T"], salt=opts["salt_concentration"], binary=binary)
    shifts = None
    if binary:
        conf_step, *_, words, shifts = io.read_conf_binary(conf_path, s.N, with_shift=True)
        if reload_path and not ckpt_reload:
            opts["seed"] = words[0] | (words[1] << 16) | (words[2] << 32)
    else:
        conf_step, _, *_ = io.read_conf(conf_path, s.N)
    start = 0 if restart else conf_step
    steps = _num(inp, "steps", int)
    dt = opts["dt"]
    energy_every = _num(inp, "print_energy_every", int, 0)
    conf_every = _num(inp, "print_conf_interval", int, 0)
    fix_every = _num(inp, "fix_diffusion_every", int, 100000) if _bool(inp.get("fix_diffusion", "true")) else 0
    device = _num(inp, "CUDA_device", int, 0)
    energy_path = path(inp.get("energy_file", "energy.dat"))
    traj_path = path(inp["trajectory_file"]) if "trajectory_file" in inp else None
    last_path = path(inp.get("lastconf_file", "last_conf.dat"))
    ckpt_every = _num(inp, "checkpoint_every", int, 0)
    ckpt_path = path(inp["checkpoint_file"]) if "checkpoint_file" in inp else None
    back_in_box = _bool(inp.get("back_in_box", "false"))
    ext = None
    if _bool(inp.get("external_forces", "false")):
        ext = io.read_external_forces(path(inp["external_forces_file"]), s,
                                      as_json=_bool(inp.get("external_forces_as_JSON", "false")))
    reload = None
    if ckpt_reload:
        with open(reload_path, "rb") as f:
            head = json.loads(f.readline())
            reload = f.read()
        start = int(head["step"])

    lines = []
    with HIPBatch(s, device=device, **opts) as b:
        b.set_sort_interval(_num(inp, "CUDA_sort_every", int, 0))
        b.set_fix_diffusion(0)   # re-boxing is done here, at the reference's steps
        if shifts is not None and shifts.any():
            b.set_images(shifts)
        if ext is not None:
            # the forces of the run's first step read the step counter `start`
            b.set_external_forces(ext, step0=start)
        if reload is not None:
            b.restore(reload)

        def snapshot(step):
            """(U, K, state with stored positions, state as printed)."""
            E = b.energies()[0]
            U, K = E[0] / s.N, (E[2] + E[3]) / s.N
            st = b.state(0)
            t = s.copy()
            t.pos, t.a1, t.a3, t.vel, t.L = st["pos"], st["a1"], st["a3"], st["vel"], st["L"]
            t.box = b.box(0)   # the barostat changes it
            img = b.images(0)
            shown = t.copy()
            if back_in_box:
                # each strand by the whole boxes that bring its absolute
                # centre of mass into the box
                absolute = t.pos + img * t.box
                for k in np.unique(s.strand):
                    m = s.strand == k
                    shown.pos[m] = t.pos[m] - np.floor(absolute[m].mean(axis=0) / t.box) * t.box
            else:
                shown.pos = t.pos + img * t.box
            return U, K, t, shown, img

        energy_f = open(energy_path, "w")
        traj_f = open(traj_path, "w") if traj_path else None
        try:
            cur = start
            plan = output_steps(start, steps, energy_every, conf_every)
            fixes = [t for t in range(start + fix_every, start + steps + 1, fix_every)] if fix_every > 0 else []
            ckpts = ([t for t in range(start + ckpt_every, start + steps + 1, ckpt_every)]
                     if ckpt_every > 0 and ckpt_path else [])
            blocks = [k for k in inp if k.startswith("data_output_") and k.rsplit("_", 1)[1].isdigit()]
            outputs = [DataOutput(inp[k], dt, path, inp, s.N) for k in sorted(blocks, key=lambda x: int(x.rsplit("_", 1)[1]))]
            out_steps = [set(o.steps(start, start + steps)) for o in outputs]
            baro = barostat_plan(inp, opts["seed"], start, start + steps)
            events = sorted(set([t for t, _, _ in plan] + fixes + ckpts + [t for o in out_steps for t in o]
                                + list(baro["steps"])))
            todo = {t: (e, c) for t, e, c in plan}
            end = start + steps
            for t in events:
                if t > cur:
                    b.run(t - cur)
                    cur = t
                if t in baro["steps"]:
                    acc, _ = b.barostat_move(baro["P"], baro["delta_L"], baro["rng"].random(4), baro["isotropic"],
                                             baro["molecular"])
                    baro["attempts"] += 1
                    baro["accepted"] += int(acc)
                e, c = todo.get(t, (False, False))
                fix = t in fixes
                # inside the loop the reference prints a due configuration,
                # then re-boxes, then prints the observables; after the last
                # step it re-boxes first
                if fix and t == end:
                    b.fix_diffusion()
                state = None
                if c and traj_f:
                    U, K, _, shown, _ = snapshot(t)
                    traj_f.write(_conf_text(shown, t, U, K))
                    traj_f.flush()
                if fix and t < end:
                    b.fix_diffusion()
                if e:
                    U, K, *_ = snapshot(t)
                    line = energy_line(t, dt, U, K)
                    energy_f.write(line + "\n")
                    energy_f.flush()
                    lines.append(line)
                for o, ts in zip(outputs, out_steps):
                    if t in ts:
                        o.f.write(o.line(t, b, s.N) + "\n")
                        o.f.flush()
                if t in ckpts:
                    data = b.checkpoint()
                    with open(ckpt_path, "wb") as f:
                        f.write((json.dumps({"step": t}) + "\n").encode())
                        f.write(data)
            if cur < start + steps:
                b.run(start + steps - cur)
                cur = start + steps
            U, K, stored, state, img = snapshot(cur)
            io.write_conf(last_path, state, step=cur, energies=(U + K, U, K))
            if "lastconf_file_bin" in inp:
                io.write_conf_binary(path(inp["lastconf_file_bin"]), stored, step=cur, energies=(U + K, U, K), shift=img)
        finally:
            energy_f.close()
            if traj_f:
                traj_f.close()
            for o in locals().get("outputs", []):
                o.close()
    return dict(energy_lines=lines, final=state, step=cur, U=U, K=K,
                barostat=dict(attempts=baro["attempts"], accepted=baro["accepted"]) if baro["attempts"] else None)


def _conf_text(s, step, U, K) -> str:
    rows = [f"t = {step}", "b = {:.17g} {:.17g} {:.17g}".format(*s.box), f"E = {U + K:.15g} {U:.15g} {K:.15g}"]
    for i in range(s.N):
        r = list(s.pos[i]) + list(s.a1[i]) + list(s.a3[i]) + list(s.vel[i]) + list(s.L[i])
        rows.append(" ".join(f"{x:.17g}" for x in r))
    return "\n".join(rows) + "\n"


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv:
        print(__doc__.split("\n\n")[1], file=sys.stderr)
        return 2
    res = run(argv[0], argv[1:])
    print(f"{res['step']} steps done: U/N {res['U']:.6f} K/N {res['K']:.6f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
