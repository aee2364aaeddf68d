"""oxDNA on-disk formats: input files, topologies and configurations.

Formats as the reference reads and writes them:
  - input file: `key = value` lines, `#` comments, `{ ... }` blocks kept as
    raw text (src/Utilities/parse_input/parse_input.cpp);
  - topology, old style: header "N N_strands", then per particle
    "strand base n3 n5"; new style: header "N N_strands 5->3", then one
    strand per line listed 5'->3' (src/Utilities/TopologyParser.cpp:15-172,
    src/Interactions/DNAInteraction.cpp:1506-1577);
  - configuration: "t = ", "b = Lx Ly Lz", "E = ..." headers, then
    "x y z a1 a3 [v L]" per particle; a1 and a3 are re-orthonormalised on
    read and a2 = a3 x a1 (src/Backends/SimBackend.cpp:605-623,
    src/Observables/Configurations/Configuration.cpp:85-140).
"""
from __future__ import annotations

import re

import numpy as np

from .systems import System

# src/Utilities/Utils.cpp:20-37 (U is read as T, N_T = 3)
_BASES = {"A": 0, "G": 1, "C": 2, "T": 3, "U": 3, "a": 0, "g": 1, "c": 2, "t": 3, "u": 3}
_NAMES = "AGCT"


def base_name(btype: int) -> str:
    if 0 <= btype < 4:
        return _NAMES[btype]
    return str(btype)


def decode_base(tok: str) -> int:
    if tok in _BASES:
        return _BASES[tok]
    return int(tok)


def parse_input(text: str) -> dict:
    """key = value pairs of an oxDNA input file.  Blocks in braces are kept
    verbatim as the value string.  Later keys override earlier ones."""
    out: dict = {}
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        line = lines[i].split("#", 1)[0].strip()
        i += 1
        if not line or "=" not in line:
            continue
        key, val = [x.strip() for x in line.split("=", 1)]
        if val.startswith("{") and val.count("{") > val.count("}"):
            depth = val.count("{") - val.count("}")
            block = [val]
            while depth > 0 and i < len(lines):
                block.append(lines[i])
                depth += lines[i].count("{") - lines[i].count("}")
                i += 1
            val = "\n".join(block)
        out[key] = val
    return out


def read_input(path: str, overrides=()) -> dict:
    with open(path) as f:
        d = parse_input(f.read())
    for ov in overrides:
        k, v = ov.split("=", 1)
        d[k.strip()] = v.strip()
    return d


def temperature(raw: str) -> float:
    """T in simulation units from "20C", "300K" or a plain number
    (src/Utilities/Utils.cpp:316-352)."""
    m = re.match(r"^\s*([-+0-9.eE]+)\s*([cCkK]?)\s*$", str(raw))
    if not m:
        raise ValueError(f"Unrecognizable temperature '{raw}'")
    v, unit = float(m.group(1)), m.group(2).lower()
    if unit == "c":
        return (v + 273.15) * 0.1 / 300.0
    if unit == "k":
        return v * 0.1 / 300.0
    return v


def write_topology(path: str, s: System, new_style: bool = False) -> None:
    N = s.N
    ns = int(s.strand.max()) + 1 if N else 0
    with open(path, "w") as f:
        if not new_style:
            f.write(f"{N} {ns}\n")
            for i in range(N):
                f.write(f"{s.strand[i] + 1} {base_name(int(s.btype[i]))} {s.n3[i]} {s.n5[i]}\n")
            return
        f.write(f"{N} {ns} 5->3\n")
        for k in range(ns):
            idx = np.nonzero(s.strand == k)[0]
            # 5' end: the particle without an n5 inside the strand
            head = [i for i in idx if s.n5[i] < 0]
            circular = not head
            i = int(head[0]) if head else int(idx[0])
            seq = []
            for _ in range(len(idx)):
                b = int(s.btype[i])
                seq.append(_NAMES[b] if 0 <= b < 4 else f"({b})")
                i = int(s.n3[i])
                if i < 0:
                    break
            f.write("".join(seq) + (" circular=true" if circular else "") + "\n")


def read_topology(path: str):
    """Returns (btype, n3, n5, strand) arrays for either topology style."""
    with open(path) as f:
        lines = [l.rstrip("\n") for l in f]
    head = lines[0].split()
    N = int(head[0])
    btype = np.zeros(N, np.int32)
    n3 = -np.ones(N, np.int32)
    n5 = -np.ones(N, np.int32)
    strand = np.zeros(N, np.int32)
    if len(head) == 3 and head[2] == "5->3":
        idx = 0
        for k, line in enumerate(l for l in lines[1:] if l.strip()):
            toks = line.split()
            seq = toks[0]
            circular = any(t.lower() in ("circular=true", "circular=1") for t in toks[1:])
            bts = []
            j = 0
            while j < len(seq):
                if seq[j] == "(":
                    e = seq.index(")", j)
                    bts.append(int(seq[j + 1:e]))
                    j = e + 1
                else:
                    bts.append(decode_base(seq[j]))
                    j += 1
            n = len(bts)
            for m in range(n):
                i = idx + m
                btype[i] = bts[m]
                strand[i] = k
                n5[i] = i - 1 if m > 0 else (idx + n - 1 if circular else -1)
                n3[i] = i + 1 if m < n - 1 else (idx if circular else -1)
            idx += n
        if idx != N:
            raise ValueError(f"topology header says {N} particles, found {idx}")
        return btype, n3, n5, strand
    i = 0
    for line in lines[1:]:
        if not line.strip() or line.startswith("#"):
            continue
        st, b, a, c = line.split()[:4]
        strand[i] = int(st) - 1
        btype[i] = decode_base(b)
        n3[i] = int(a)
        n5[i] = int(c)
        i += 1
    if i != N:
        raise ValueError(f"topology header says {N} particles, found {i}")
    return btype, n3, n5, strand


def write_conf(path: str, s: System, step: int = 0, energies=(0.0, 0.0, 0.0)) -> None:
    vel = s.vel if s.vel is not None else np.zeros((s.N, 3))
    L = s.L if s.L is not None else np.zeros((s.N, 3))
    with open(path, "w") as f:
        f.write(f"t = {step}\n")
        f.write("b = {:.17g} {:.17g} {:.17g}\n".format(*s.box))
        f.write("E = {:.15g} {:.15g} {:.15g}\n".format(*energies))
        for i in range(s.N):
            row = list(s.pos[i]) + list(s.a1[i]) + list(s.a3[i]) + list(vel[i]) + list(L[i])
            f.write(" ".join(f"{x:.17g}" for x in row) + "\n")


def read_conf(path: str, N: int):
    """Returns (step, box, pos, a1, a2, a3, vel, L) with the reference's
    re-orthonormalisation of a1 and a3 (src/Backends/SimBackend.cpp:616-623)."""
    with open(path) as f:
        t = f.readline()
        b = f.readline()
        f.readline()
        step = int(float(t.split("=")[1]))
        box = np.array([float(x) for x in b.split("=")[1].split()])
        data = np.loadtxt(f, ndmin=2)
    if data.shape[0] != N:
        raise ValueError(f"configuration has {data.shape[0]} particles, expected {N}")
    pos = data[:, 0:3].copy()
    a1 = data[:, 3:6] / np.linalg.norm(data[:, 3:6], axis=1, keepdims=True)
    a3 = data[:, 6:9] / np.linalg.norm(data[:, 6:9], axis=1, keepdims=True)
    a1 = a1 - a3 * np.sum(a1 * a3, axis=1, keepdims=True)
    a1 /= np.linalg.norm(a1, axis=1, keepdims=True)
    a2 = np.cross(a3, a1)
    a2 /= np.linalg.norm(a2, axis=1, keepdims=True)
    vel = data[:, 9:12].copy() if data.shape[1] >= 15 else np.zeros((N, 3))
    L = data[:, 12:15].copy() if data.shape[1] >= 15 else np.zeros((N, 3))
    return step, box, pos, a1, a2, a3, vel, L


# the reference's binary configuration (src/Observables/Configurations/
# BinaryConfiguration.cpp:17-91, read by src/Backends/SimBackend.cpp:521-553,
# 636-648), little-endian: the step (int64), the drand48 seed words (3 x
# uint16), the box (3 doubles), total, potential and kinetic energy per
# nucleotide (3 doubles), then per nucleotide the position (3 doubles), its
# image shift (3 int32), the rows of the orientation's transpose a1, a2, a3
# (9 doubles), velocity and angular momentum (6 doubles); the observable
# output ends the configuration with a newline
_BIN_HEAD = np.dtype([("step", "<i8"), ("seed", "<u2", 3), ("box", "<f8", 3), ("E", "<f8", 3)])
_BIN_PART = np.dtype([("pos", "<f8", 3), ("shift", "<i4", 3), ("a", "<f8", 9), ("vel", "<f8", 3), ("L", "<f8", 3)])


def write_conf_binary(path: str, s: System, step: int = 0, energies=(0.0, 0.0, 0.0), seed=(0, 0, 0), shift=None) -> None:
    """Writes the reference's binary configuration.  `energies` = (total,
    potential, kinetic) per nucleotide; `seed` = the three drand48 state
    words the reference stores (its generator's state; this library draws
    from Philox, so they carry nothing here).  Positions are the stored
    ones (moved by whole boxes when strands are re-boxed) and `shift` the
    (N, 3) periodic image counts, as the reference writes p->pos and its
    _pos_shift (src/Observables/Configurations/BinaryConfiguration.cpp:
    55-70); None writes zeros."""
    vel = s.vel if s.vel is not None else np.zeros((s.N, 3))
    L = s.L if s.L is not None else np.zeros((s.N, 3))
    head = np.zeros(1, _BIN_HEAD)
    head["step"], head["seed"], head["box"], head["E"] = step, seed, s.box, energies
    rec = np.zeros(s.N, _BIN_PART)
    rec["pos"] = s.pos
    if shift is not None:
        rec["shift"] = shift
    rec["a"] = np.concatenate([s.a1, np.cross(s.a3, s.a1), s.a3], axis=1)
    rec["vel"] = vel
    rec["L"] = L
    with open(path, "wb") as f:
        f.write(head.tobytes())
        f.write(rec.tobytes())
        f.write(b"\n")


def read_conf_binary(path: str, N: int, with_shift: bool = False):
    """Returns (step, box, pos, a1, a2, a3, vel, L, seed) of the reference's
    binary configuration (and the (N, 3) image shifts last with
    with_shift); the axes are taken as stored (the reference does not
    re-orthonormalise them here)."""
    with open(path, "rb") as f:
        raw = f.read()
    need = _BIN_HEAD.itemsize + N * _BIN_PART.itemsize
    if len(raw) < need:
        raise ValueError(f"binary configuration of {len(raw)} bytes, {need} expected for {N} particles")
    head = np.frombuffer(raw, _BIN_HEAD, 1)[0]
    rec = np.frombuffer(raw, _BIN_PART, N, _BIN_HEAD.itemsize)
    a = rec["a"]
    out = (int(head["step"]), head["box"].copy(), rec["pos"].copy(), a[:, 0:3].copy(), a[:, 3:6].copy(), a[:, 6:9].copy(),
           rec["vel"].copy(), rec["L"].copy(), tuple(int(x) for x in head["seed"]))
    return out + (rec["shift"].astype(np.int32),) if with_shift else out


def load_system(topology: str, conf: str, T: float = 0.1, salt: float = 0.5, name: str = "", binary: bool = False) -> System:
    btype, n3, n5, strand = read_topology(topology)
    if binary:
        step, box, pos, a1, a2, a3, vel, L, _ = read_conf_binary(conf, len(btype))
    else:
        step, box, pos, a1, a2, a3, vel, L = read_conf(conf, len(btype))
    return System(pos=pos, a1=a1, a3=a3, btype=btype, n3=n3, n5=n5, strand=strand, box=box, T=T, salt=salt,
                  name=name, vel=vel, L=L)


EXT_TYPES = ("string", "trap", "mutual_trap", "constant_trap", "repulsion_plane", "sphere", "lowdim_trap",
             "attraction_plane", "repulsion_plane_moving", "sphere_smooth", "repulsive_sphere_moving",
             "repulsive_kepler_poinsot", "LJ_wall", "twist", "generic_central_force", "LJ_cone", "ellipsoid", "com",
             "yukawa_sphere")


def _vec3(v) -> tuple:
    if isinstance(v, (list, tuple)):
        return tuple(float(x) for x in v)
    parts = str(v).replace(" ", "").split(",")
    if len(parts) != 3:
        raise ValueError(f"could not parse the vector {v!r}")
    return tuple(float(x) for x in parts)


def particles_from_string(text, s: System) -> list:
    """Particle indices of an external force's `particle` key, as the
    reference's Utils::get_particles_from_string (src/Utilities/Utils.cpp:
    409-500): comma-separated indices, `last`, `all` or -1 (every
    particle), and ranges a-b that walk from a along the 5' neighbours (n5)
    until b, both included."""
    out = []
    for tok in str(text).replace(" ", "").split(","):
        if "-" in tok and not tok.startswith("-"):
            a, b = tok.split("-", 1)
            p = [s.N - 1 if x == "last" else int(x) for x in (a, b)]
            for x in p:
                if not 0 <= x < s.N:
                    raise ValueError(f"particle {x} out of range")
            j = p[0]
            while True:
                out.append(j)
                if j == p[1]:
                    break
                j = int(s.n5[j])
                if j < 0 or j == p[0]:
                    raise ValueError(f"could not get from particle {p[0]} to particle {p[1]}")
        elif tok == "last":
            out.append(s.N - 1)
        elif tok in ("all", "-1"):
            out.extend(range(s.N))
        else:
            j = int(tok)
            if not 0 <= j < s.N:
                raise ValueError(f"particle {j} out of range")
            out.append(j)
    return out


def _blocks(text: str, what: str) -> list:
    """key = value blocks in braces, as the reference's external_forces_file
    and order_parameters_file hold them (src/Forces/ForceFactory.cpp:
    140-200, src/Utilities/OrderParameters.cpp:301-340: '#' comments, one
    entry per {...})."""
    clean = "\n".join(line.split("#", 1)[0] for line in text.splitlines())
    blocks, depth, cur = [], 0, []
    for ch in clean:
        if ch == "{":
            depth += 1
            if depth > 1:
                raise ValueError(f"{what}: nested braces")
            cur = []
        elif ch == "}":
            depth -= 1
            if depth < 0:
                raise ValueError(f"{what}: unbalanced braces")
            body = "".join(cur).strip()
            if not body:
                raise ValueError(f"{what}: nothing between braces")
            blocks.append(parse_input(body))
        elif depth == 1:
            cur.append(ch)
    if depth != 0:
        raise ValueError(f"{what}: unbalanced braces")
    return blocks


def _force_blocks(text: str) -> list:
    return _blocks(text, "external forces file")


def _flag(v) -> bool:
    return str(v).strip().lower() in ("1", "true", "yes")


def _ext_record(t: str, b: dict, s: System):
    """One block of an external_forces_file as the reference's force class
    reads it in its init (src/Forces/<Class>.cpp), defaults and key
    handling included: the record's keys are the fields of the C record
    (include/oxh.h oxh_ext_force).  Returns (record, particle indices)."""
    def need(*keys):
        for k in keys:
            if k not in b:
                raise ValueError(f"external force {t}: missing key {k}")

    def num(k, default=None):
        if k in b:
            return float(b[k])
        if default is None:
            raise ValueError(f"external force {t}: missing key {k}")
        return default

    def vec(k, default=None):
        if k in b:
            return _vec3(b[k])
        if default is None:
            raise ValueError(f"external force {t}: missing key {k}")
        return default

    rec = {"type": t}
    if t == "string":                      # ConstantRateForce.cpp:20-42
        rec.update(F0=num("F0"), rate=num("rate"), dir=vec("dir"), dir_as_centre=_flag(b.get("dir_as_centre", 0)))
    elif t == "trap":                      # MovingTrap.cpp:21-53
        rec.update(stiff=num("stiff"), rate=num("rate"), dir=vec("dir"), pos0=vec("pos0"))
    elif t in ("mutual_trap", "constant_trap"):   # MutualTrap.cpp:24-60, ConstantTrap.cpp:22-47
        rec.update(stiff=num("stiff"), r0=num("r0"), PBC=_flag(b.get("PBC", 0)), ref_particle=int(num("ref_particle")))
        if t == "mutual_trap":
            rec.update(rate=num("rate", 0.0), stiff_rate=num("stiff_rate", 0.0))
        need("particle")
        pid = int(b["particle"])
        if pid == -1:
            raise ValueError(f"{t} cannot act on every particle")
        return rec, [pid]
    elif t == "repulsion_plane":           # RepulsionPlane.cpp:22-54
        rec.update(stiff=num("stiff"), position=num("position"), dir=vec("dir"), v=num("v", 0.0),
                   end_position=num("end_position", 1e6))
    elif t == "sphere":                    # RepulsiveSphere.cpp:22-48
        rec.update(stiff=num("stiff"), r0=num("r0"), rate=num("rate", 0.0), r_ext=num("r_ext", 1e10),
                   center=vec("center", (0.0, 0.0, 0.0)))
    elif t == "lowdim_trap":               # LowdimMovingTrap.cpp:21-63
        need("visibility")
        vis = [int(x) for x in str(b["visibility"]).replace(" ", "").split(",")]
        if len(vis) != 3:
            raise ValueError(f"external force {t}: could not parse visibility {b['visibility']!r}")
        rec.update(stiff=num("stiff"), rate=num("rate"), pos0=vec("pos0"), dir=vec("dir"), visibility=tuple(vis))
    elif t == "attraction_plane":          # AttractionPlane.cpp:20-41
        rec.update(stiff=num("stiff"), position=num("position"), dir=vec("dir"))
    elif t == "repulsion_plane_moving":    # RepulsionPlaneMoving.cpp:17-46
        need("ref_particle")
        refs = sorted(particles_from_string(b["ref_particle"], s))
        if refs[-1] - refs[0] + 1 != len(refs):
            raise ValueError("repulsion_plane_moving needs a contiguous list of ref_particle indices")
        rec.update(stiff=num("stiff"), dir=vec("dir"), ref_particle=refs[0], ref_last=refs[-1])
    elif t == "sphere_smooth":             # RepulsiveSphereSmooth.cpp:22-48
        # the reference reads r_ext into smooth and alpha as well
        # (RepulsiveSphereSmooth.cpp:30-32), so r_ext is required
        r_ext = num("r_ext")
        rec.update(stiff=num("stiff"), r0=num("r0"), r_ext=r_ext, smooth=r_ext, alpha=r_ext,
                   center=vec("center", (0.0, 0.0, 0.0)))
    elif t == "repulsive_sphere_moving":   # RepulsiveSphereMoving.cpp:24-85
        center = vec("center", (0.0, 0.0, 0.0))
        steps = int(num("steps", num("move_steps", 0.0)))
        rec.update(stiff=num("stiff"), r0=num("r0"), rate=num("rate", 0.0), r_ext=num("r_ext", 1e10),
                   pos0=vec("origin", center), target=vec("target", (0.0, 0.0, 0.0)), steps=steps)
    elif t == "repulsive_kepler_poinsot":  # RepulsiveKeplerPoinsot.cpp:34-103
        rec.update(stiff=num("stiff"), rate=num("rate", 0.0), apex=num("apex", 1.2), base=num("base", 0.7),
                   base_radius=num("base_radius", 0.45), kappa=num("kappa", 25.0), center=vec("center", (0.0, 0.0, 0.0)))
        if rec["apex"] <= rec["base"]:
            raise ValueError("repulsive_kepler_poinsot: require apex > base")
        if rec["base_radius"] <= 0:
            raise ValueError("repulsive_kepler_poinsot: require base_radius > 0")
        if rec["kappa"] <= 0:
            rec["kappa"] = 25.0
    elif t in ("LJ_wall", "LJ_cone"):      # LJWall.cpp:26-57, LJCone.cpp:26-62
        n = int(num("n", 6.0))
        if n % 2 != 0:
            raise ValueError(f"{t}: n ({n}) should be an even integer")
        cutoff = 2.0 ** (1.0 / n) if _flag(b.get("only_repulsive", 0)) else 1e6
        rec.update(stiff=num("stiff", 1.0), sigma=num("sigma", 1.0), n=n, cutoff=cutoff, dir=vec("dir"))
        if t == "LJ_wall":
            rec["position"] = num("position")
        else:
            rec.update(alpha=num("alpha"), pos0=vec("pos0"))
    elif t == "twist":                     # ConstantRateTorque.cpp:17-60
        rec.update(stiff=num("stiff"), rate=num("rate"), F0=num("base"), axis=vec("axis"), pos0=vec("pos0"),
                   center=vec("center"), mask=vec("mask", (0.0, 0.0, 0.0)))
    elif t == "generic_central_force":     # GenericCentralForce.cpp:130-172
        need("force_type")
        ft = str(b["force_type"]).strip()
        if ft == "interpolated":
            raise ValueError("generic_central_force: force_type = interpolated (a potential_file table) is not supported")
        if ft != "gravity":
            raise ValueError(f"GenericCentralForce: unsupported type {ft!r}")
        rec.update(F0=num("F0"), inner_cut_off=num("inner_cut_off", 0.0), outer_cut_off=num("outer_cut_off", 0.0),
                   center=vec("center"))
    elif t == "ellipsoid":                 # RepulsiveEllipsoid.cpp:22-49
        rec.update(stiff=num("stiff"), r_2=vec("r_2"), r_1=vec("r_1", (1e-6, 1e-6, 1e-6)),
                   center=vec("center", (0.0, 0.0, 0.0)))
    elif t == "com":                       # COMForce.cpp:22-42
        need("com_list", "ref_list")
        com = particles_from_string(b["com_list"], s)
        ref = particles_from_string(b["ref_list"], s)
        rec.update(stiff=num("stiff"), r0=num("r0"), rate=num("rate", 0.0), com_list=tuple(com), ref_list=tuple(ref))
        return rec, com
    elif t == "yukawa_sphere":             # YukawaSphere.cpp:22-54
        # the reference reads WCA_n into the WCA sigma and keeps n = 6
        # (YukawaSphere.cpp:27-29)
        sigma = num("WCA_sigma", 1.0)
        if "WCA_n" in b:
            sigma = float(b["WCA_n"])
        debye_length = num("debye_length")
        rec.update(radius=num("radius"), epsilon=num("WCA_epsilon", 1.0), sigma=sigma, n=6, debye_length=debye_length,
                   debye_A=num("debye_A"), cutoff=num("cutoff", 4.0 * debye_length), center=vec("center", (0.0, 0.0, 0.0)))
    else:
        raise ValueError(f"external force type {t!r} is not supported (supported: {', '.join(EXT_TYPES)})")
    need("particle")
    return rec, particles_from_string(b["particle"], s)


def read_external_forces(path: str, s: System, as_json: bool = False) -> list:
    """The forces of an external_forces_file, expanded to one record per
    particle in the file's order: dicts whose keys are the fields of the C
    record (type, particle, F0, rate, stiff, dir, center, ...), filled as
    the reference's force classes read their keys.  Types: those the
    reference's CUDA backend accepts (MD_CUDABackend::
    _apply_external_forces_changes, src/CUDA/Backends/MD_CUDABackend.cu:
    108-220) apart from the metadynamics ones, plus constant_trap; any
    other type raises."""
    with open(path) as f:
        text = f.read()
    if as_json:
        import json
        blocks = [{k: (v if isinstance(v, (list, tuple)) else str(v)) for k, v in b.items()} for b in json.loads(text)]
    else:
        blocks = _force_blocks(text)
    out = []
    for b in blocks:
        t = b.get("type")
        if t not in EXT_TYPES:
            raise ValueError(f"external force type {t!r} is not supported (supported: {', '.join(EXT_TYPES)})")
        rec, pids = _ext_record(t, b, s)
        for pid in pids:
            r = dict(rec)
            r["particle"] = pid
            out.append(r)
    return out


# ------------------------------------------------------- order parameters
HB_CUTOFF = -0.1   # src/defs.h:62


def read_order_parameters(path: str, N: int) -> list:
    """The parameters of an order_parameters_file (the reference's op_file,
    parsed as OrderParameters::init_from_file does, src/Utilities/
    OrderParameters.cpp:301-598): one {...} block per parameter with
    `order_parameter` = bond or mindistance, a `name` and keys pair1,
    pair2, ... of the form "a, b" (one pair) or "a1-a2, b1-b2" (a1 with b2,
    a1 + 1 with b2 - 1, ..., the two ranges of equal length).  Pairs are
    kept as (min, max); a bond parameter keeps each pair once.  Optional:
    cutoff (bond; the reference uses it in FFS runs only), interfaces and
    use_COM (mindistance: centre-centre instead of base-base distances).
    Returns dicts with type, name, pairs, cutoff, use_COM, interfaces.  The
    twist parameter is not supported and raises."""
    with open(path) as f:
        text = f.read()
    out = []
    for d in _blocks(text, "order parameters file"):
        t = d.get("order_parameter")
        if t is None:
            raise KeyError("order parameter without order_parameter")
        if t not in ("bond", "mindistance"):
            raise ValueError(f"order parameter type {t!r} is not supported (supported: bond, mindistance)")
        if "name" not in d:
            raise KeyError(f"{t} order parameter without a name")
        pairs = []
        for key in (k for k in d if k.startswith("pair")):
            vals = [v.strip() for v in d[key].split(",")]
            if len(vals) != 2:
                raise ValueError(f"order parameters file: invalid value {d[key]!r} for key {key}")
            ranged = ["-" in v for v in vals]
            if ranged[0] != ranged[1]:
                raise ValueError(f"order parameters file: invalid value {d[key]!r} for key {key}")
            if ranged[0]:
                (s1, e1), (s2, e2) = ([int(x) for x in v.split("-")] for v in vals)
                if s1 - e1 != s2 - e2 or s1 > e1 or s2 > e2:
                    raise ValueError(f"order parameters file: invalid list {d[key]!r}")
                new = [(m, e2 - c) for c, m in enumerate(range(s1, e1 + 1))]
            else:
                new = [(int(vals[0]), int(vals[1]))]
            for a, b in new:
                if not (0 <= a < N and 0 <= b < N):
                    raise ValueError(f"order parameters file: particle index out of range in {d[key]!r}")
                pairs.append((min(a, b), max(a, b)))
        if t == "bond":
            pairs = sorted(set(pairs))
        interfaces = [float(x) for x in d["interfaces"].split(",")] if "interfaces" in d else []
        use_com = d.get("use_COM", "false").strip().lower() in ("1", "true", "yes")
        out.append(dict(type=t, name=d["name"], pairs=pairs, cutoff=float(d.get("cutoff", HB_CUTOFF)),
                        use_COM=use_com, interfaces=interfaces))
    return out
