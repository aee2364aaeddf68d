"""Synthetic oxDNA systems for the workloads this package is measured on.

Geometry follows the reference's B-helix builder, utils/generate-sa.py:
rise 0.3897628551303122 per base pair, twist 35.9 degrees per base pair,
nucleotide centre at base point - 0.6 * a1 (generate-sa.py:53-54,127-221),
second strand antiparallel with complementary, reversed sequence.  Strand
numbering follows the reference's old-style topology: inside a strand,
particle i has n3 = i - 1 and n5 = i + 1 (generate-sa.py:294-303).

Base types: A=0, G=1, C=2, T=3; complementary pairs sum to 3
(src/Interactions/DNAInteraction.cpp:785).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

BASE_BASE = 0.3897628551303122
CM_CENTER_DS = 0.4 + 0.2
TWIST = 35.9 * math.pi / 180.0
FENE_R0 = 0.7564
BACK1, BACK2 = -0.34, 0.3408


@dataclasses.dataclass
class System:
    """One oxDNA system: per-nucleotide state plus topology and box."""

    pos: np.ndarray      # (N, 3) float64
    a1: np.ndarray       # (N, 3) principal axis
    a3: np.ndarray       # (N, 3) stacking axis
    btype: np.ndarray    # (N,) int32
    n3: np.ndarray       # (N,) int32, -1 = none
    n5: np.ndarray       # (N,) int32, -1 = none
    strand: np.ndarray   # (N,) int32
    box: np.ndarray      # (3,) float64
    T: float = 0.1       # simulation units
    salt: float = 0.5
    name: str = ""
    vel: np.ndarray | None = None
    L: np.ndarray | None = None

    @property
    def N(self) -> int:
        return int(self.pos.shape[0])

    @property
    def a2(self) -> np.ndarray:
        return np.cross(self.a3, self.a1)

    def copy(self) -> "System":
        return dataclasses.replace(self, **{f.name: (getattr(self, f.name).copy()
                                   if isinstance(getattr(self, f.name), np.ndarray) else getattr(self, f.name))
                                   for f in dataclasses.fields(self)})


def celsius(tc: float) -> float:
    """Temperature in simulation units, as src/Utilities/Utils.cpp:326-331."""
    return (tc + 273.15) * 0.1 / 300.0


def rotation_matrix(axis, angle) -> np.ndarray:
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    c, s = math.cos(angle), math.sin(angle)
    oc = 1.0 - c
    x, y, z = axis
    return np.array([[oc * x * x + c, oc * x * y - s * z, oc * x * z + s * y],
                     [oc * x * y + s * z, oc * y * y + c, oc * y * z - s * x],
                     [oc * x * z - s * y, oc * y * z + s * x, oc * z * z + c]])


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniformly distributed rotation (normalised 4-d Gaussian quaternion)."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _helix_strand(bp, start, direction, a1_0):
    """One strand of an ideal B-helix going along +direction."""
    R = rotation_matrix(direction, TWIST)
    pos, a1s, a3s = [], [], []
    a1 = a1_0.copy()
    rb = np.array(start, dtype=np.float64)
    for i in range(bp):
        pos.append(rb - CM_CENTER_DS * a1)
        a1s.append(a1.copy())
        a3s.append(direction.copy())
        if i != bp - 1:
            a1 = R @ a1
            rb = rb + direction * BASE_BASE
    return pos, a1s, a3s, rb, a1


def duplex_geometry(bp, start=(0.0, 0.0, 0.0), direction=(0.0, 0.0, 1.0), perp=None, rng=None):
    """Positions and axes of both strands of an ideal B-DNA duplex
    (generate-sa.py:127-221).  Returns (pos, a1, a3) with 2*bp rows; the
    first bp rows are strand 1 (3'->5' along +direction), the next bp rows
    strand 2 coming back."""
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    if perp is None:
        rng = rng or np.random.default_rng(0)
        v = rng.random(3)
    else:
        v = np.asarray(perp, dtype=np.float64)
    v = v - d * np.dot(d, v)
    v /= np.linalg.norm(v)
    p1, a11, a31, rb, a1_last = _helix_strand(bp, start, d, v)
    a1 = -a1_last
    a3 = -d
    Rt = rotation_matrix(d, TWIST).T
    p2, a12, a32 = [], [], []
    for _ in range(bp):
        p2.append(rb - CM_CENTER_DS * a1)
        a12.append(a1.copy())
        a32.append(a3.copy())
        a1 = Rt @ a1
        rb = rb + a3 * BASE_BASE
    return np.array(p1 + p2), np.array(a11 + a12), np.array(a31 + a32)


class _Builder:
    """Accumulates strands into one System."""

    def __init__(self):
        self.pos, self.a1, self.a3, self.bt, self.n3, self.n5, self.strand = [], [], [], [], [], [], []
        self.nstrands = 0

    @property
    def N(self):
        return len(self.bt)

    def add_strand(self, pos, a1, a3, btypes, circular=False):
        start = self.N
        n = len(btypes)
        for k in range(n):
            self.pos.append(np.asarray(pos[k], dtype=np.float64))
            self.a1.append(np.asarray(a1[k], dtype=np.float64))
            self.a3.append(np.asarray(a3[k], dtype=np.float64))
            self.bt.append(int(btypes[k]))
            self.n3.append(start + k - 1 if k > 0 else (start + n - 1 if circular else -1))
            self.n5.append(start + k + 1 if k < n - 1 else (start if circular else -1))
            self.strand.append(self.nstrands)
        self.nstrands += 1

    def build(self, box, name, T=0.1, salt=0.5) -> System:
        a1 = np.array(self.a1)
        a3 = np.array(self.a3)
        a1 /= np.linalg.norm(a1, axis=1, keepdims=True)
        a3 /= np.linalg.norm(a3, axis=1, keepdims=True)
        return System(pos=np.array(self.pos), a1=a1, a3=a3, btype=np.array(self.bt, dtype=np.int32),
                      n3=np.array(self.n3, dtype=np.int32), n5=np.array(self.n5, dtype=np.int32),
                      strand=np.array(self.strand, dtype=np.int32), box=np.asarray(box, dtype=np.float64),
                      T=T, salt=salt, name=name)


def add_duplex(b: _Builder, seq, centre, R=None, perp=(1.0, 0.0, 0.0)):
    """Adds a duplex with sequence `seq` (strand 1), centred at `centre`,
    rotated by R (axis along R @ z)."""
    bp = len(seq)
    pos, a1, a3 = duplex_geometry(bp, start=(0.0, 0.0, 0.0), direction=(0.0, 0.0, 1.0), perp=perp)
    pos = pos - np.array([0.0, 0.0, (bp - 1) * BASE_BASE / 2.0])
    if R is not None:
        pos, a1, a3 = pos @ R.T, a1 @ R.T, a3 @ R.T
    pos = pos + np.asarray(centre, dtype=np.float64)
    seq2 = [3 - s for s in reversed(list(seq))]
    b.add_strand(pos[:bp], a1[:bp], a3[:bp], seq)
    b.add_strand(pos[bp:], a1[bp:], a3[bp:], seq2)


# oxRNA A-form duplex.  Helical step (twist, rise) and the pose of base
# pair 0 in the helix frame (axis z, strand 1 running 5' -> 3' along +z):
# centre, a1 and a3 of the strand-1 and strand-2 nucleotides, averaged over
# the base pairs of a relaxed 8-bp oxRNA duplex (the reference's
# examples/RNA_DUPLEX_MELT/init.conf) after undoing each pair's step.
RNA_TWIST = 0.6000194497435797
RNA_RISE = 0.3614667206242939
RNA_BP = (((0.789424, -0.101093, -0.0306), (-0.639166, 0.732138, -0.235459), (0.103122, -0.221809, -0.969622)),
          ((0.050729, 0.795566, -0.313699), (0.645737, -0.73524, 0.206025), (-0.225051, 0.074571, 0.971489)))


def rna_duplex_geometry(bp: int):
    """Centres, a1 and a3 of an oxRNA duplex of bp base pairs along z,
    centred at the origin, in the builder's strand order (3' end first):
    rows [0, bp) strand 1, [bp, 2 bp) strand 2; base pair k of strand 1
    (5' -> 3') pairs with row bp + k."""
    out = {1: [], 2: []}
    z0 = (bp - 1) * RNA_RISE / 2.0
    for k in range(bp):
        R = rotation_matrix((0.0, 0.0, 1.0), k * RNA_TWIST)
        for strand in (1, 2):
            c, a1, a3 = (np.array(v, dtype=np.float64) for v in RNA_BP[strand - 1])
            a1 = a1 / np.linalg.norm(a1)
            a3 = a3 - (a3 @ a1) * a1
            a3 = a3 / np.linalg.norm(a3)
            out[strand].append((R @ c + np.array([0.0, 0.0, k * RNA_RISE - z0]), R @ a1, R @ a3))
    rows = list(reversed(out[1])) + out[2]
    return (np.array([r[0] for r in rows]), np.array([r[1] for r in rows]), np.array([r[2] for r in rows]))


def add_rna_duplex(b: _Builder, seq, centre, R=None):
    """Adds an oxRNA duplex whose strand 1 reads `seq` 5' -> 3' (A=0, G=1,
    C=2, U=3), centred at `centre`, axis along R @ z."""
    bp = len(seq)
    pos, a1, a3 = rna_duplex_geometry(bp)
    if R is not None:
        pos, a1, a3 = pos @ R.T, a1 @ R.T, a3 @ R.T
    pos = pos + np.asarray(centre, dtype=np.float64)
    b.add_strand(pos[:bp], a1[:bp], a3[:bp], list(reversed(list(seq))))
    b.add_strand(pos[bp:], a1[bp:], a3[bp:], [3 - x for x in seq])


def rna_duplex_set(seed: int = 0, n_sys: int = 24, per_side: int = 4, bp: int = 32, box: float = 20.0,
                   tc: float = 37.0, salt: float = 1.0):
    """n_sys independent oxRNA systems of per_side^2 random-sequence
    duplexes of bp base pairs each (per_side^2 * 2 bp nucleotides), axes
    along z on a square lattice filling a cubic box of side `box`."""
    rng = np.random.default_rng(seed)
    out = []
    a = box / per_side
    for k in range(n_sys):
        b = _Builder()
        for ix in range(per_side):
            for iy in range(per_side):
                R = rotation_matrix((0.0, 0.0, 1.0), rng.uniform(0.0, 2.0 * math.pi))
                add_rna_duplex(b, rng.integers(0, 4, bp), ((ix + 0.5) * a, (iy + 0.5) * a, box / 2.0), R)
        out.append(b.build([box, box, box], f"rna_set{k}", T=celsius(tc), salt=salt))
    return out


def _loop_arc(B1, B2, up, n_loop, bond=FENE_R0):
    """Backbone-site positions of n_loop nucleotides on a circular arc from
    B1 to B2, bulging along `up`, consecutive sites `bond` apart."""
    c = np.linalg.norm(B2 - B1)
    m = n_loop + 1
    ratio = c / bond
    lo, hi = 1e-9, math.pi / m - 1e-12
    for _ in range(200):   # sin(m x)/sin(x) decreases from m to 0 on (0, pi/m)
        mid = 0.5 * (lo + hi)
        if math.sin(m * mid) / math.sin(mid) > ratio:
            lo = mid
        else:
            hi = mid
    x = 0.5 * (lo + hi)
    R = bond / (2.0 * math.sin(x))
    half = m * x                       # half of the subtended angle
    u = (B2 - B1) / c
    w = up - u * np.dot(up, u)
    w /= np.linalg.norm(w)
    mid = 0.5 * (B1 + B2)
    # centre sits on the side of the bulge when the arc is the major one
    h = R * math.cos(half)
    centre = mid - w * h
    pts = []
    e1 = (B1 - centre) / R
    e2 = np.cross(np.cross(e1, w), e1)
    e2 = w - e1 * np.dot(w, e1)
    e2 /= np.linalg.norm(e2)
    for k in range(1, m):
        ang = 2.0 * x * k
        pts.append(centre + R * (math.cos(ang) * e1 + math.sin(ang) * e2))
    return np.array(pts), centre


def add_hairpin(b: _Builder, stem_seq, loop_seq, centre, R=None):
    """A hairpin: stem as an ideal B-helix, loop nucleotides on an arc that
    joins the two stem strands with FENE-equilibrium backbone spacing, bases
    pointing away from the arc's centre."""
    s = len(stem_seq)
    pos, a1, a3 = duplex_geometry(s, direction=(0.0, 0.0, 1.0), perp=(1.0, 0.0, 0.0))
    back = pos + BACK1 * a1 + BACK2 * np.cross(a3, a1)
    top1, top2 = s - 1, s
    nl = len(loop_seq)
    up = np.array([0.0, 0.0, 1.0])
    pts, cen = _loop_arc(back[top1], back[top2], up, nl)
    lpos, la1, la3 = [], [], []
    prev = back[top1]
    for k in range(nl):
        B = pts[k]
        nxt = pts[k + 1] if k + 1 < nl else back[top2]
        t = nxt - prev
        t /= np.linalg.norm(t)
        inward = cen - B
        inward -= t * np.dot(inward, t)
        inward /= np.linalg.norm(inward)
        aa1 = -inward   # bases point out of the loop, clear of their neighbours
        aa3 = -t
        aa2 = np.cross(aa3, aa1)
        lpos.append(B - (BACK1 * aa1 + BACK2 * aa2))
        la1.append(aa1)
        la3.append(aa3)
        prev = B
    stem2 = [3 - x for x in reversed(list(stem_seq))]
    allpos = np.concatenate([pos[:s], np.array(lpos).reshape(-1, 3), pos[s:]])
    alla1 = np.concatenate([a1[:s], np.array(la1).reshape(-1, 3), a1[s:]])
    alla3 = np.concatenate([a3[:s], np.array(la3).reshape(-1, 3), a3[s:]])
    allpos = allpos - allpos.mean(axis=0)
    if R is not None:
        allpos, alla1, alla3 = allpos @ R.T, alla1 @ R.T, alla3 @ R.T
    allpos = allpos + np.asarray(centre)
    b.add_strand(allpos, alla1, alla3, list(stem_seq) + list(loop_seq) + stem2)


def add_single_strand(b: _Builder, seq, centre, R=None):
    """A single strand laid out as one strand of an ideal helix."""
    n = len(seq)
    pos, a1, a3, _, _ = _helix_strand(n, (0.0, 0.0, 0.0), np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0]))
    pos = np.array(pos)
    pos -= pos.mean(axis=0)
    a1, a3 = np.array(a1), np.array(a3)
    if R is not None:
        pos, a1, a3 = pos @ R.T, a1 @ R.T, a3 @ R.T
    b.add_strand(pos + np.asarray(centre), a1, a3, seq)


# ---------------------------------------------------------------- workloads

def duplex_20bp(seed: int = 0) -> System:
    """One 20-bp duplex, random ACGT sequence, cubic box of side 20,
    T = 20 C, 0.5 M (the reference's CPU-runnable case)."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    add_duplex(b, rng.integers(0, 4, 20), (10.0, 10.0, 10.0))
    return b.build([20.0, 20.0, 20.0], "duplex20", T=celsius(20.0), salt=0.5)


def duplex_gas(seed: int = 0, n_duplex: int = 512, bp: int = 24, density: float = 1e-3, tc: float = 25.0) -> System:
    """n_duplex random-sequence duplexes, centres on a cubic lattice,
    uniformly random orientations, nucleotide density `density`."""
    rng = np.random.default_rng(seed)
    N = n_duplex * 2 * bp
    L = (N / density) ** (1.0 / 3.0)
    side = int(round(n_duplex ** (1.0 / 3.0)))
    assert side ** 3 == n_duplex
    a = L / side
    b = _Builder()
    for ix in range(side):
        for iy in range(side):
            for iz in range(side):
                c = np.array([ix + 0.5, iy + 0.5, iz + 0.5]) * a
                add_duplex(b, rng.integers(0, 4, bp), c, random_rotation(rng))
    return b.build([L, L, L], "duplex_gas", T=celsius(tc), salt=0.5)


def duplex_bundle(seed: int = 0, nx: int = 32, ny: int = 32, bp: int = 96, spacing: float = 2.6,
                  lz: float = 40.0, tc: float = 30.0, salt: float = 0.2) -> System:
    """nx*ny random-sequence duplexes, axes along z on a hexagonal lattice
    with axis spacing `spacing`, periodic orthogonal box."""
    rng = np.random.default_rng(seed)
    dy = spacing * math.sqrt(3.0) / 2.0
    Lx, Ly = nx * spacing, ny * dy
    b = _Builder()
    for j in range(ny):
        for i in range(nx):
            c = np.array([(i + 0.5 * (j % 2) + 0.25) * spacing, (j + 0.5) * dy, lz / 2.0])
            add_duplex(b, rng.integers(0, 4, bp), c, None, perp=(1.0, 0.0, 0.0))
    return b.build([Lx, Ly, lz], "duplex_bundle", T=celsius(tc), salt=salt)


def hairpin_set(seed: int = 0, n: int = 256, box: float = 30.0, t_lo: float = 25.0, t_hi: float = 75.0):
    """n independent hairpins of uneven size: stem 6-12 bp, loop 4-36 nt,
    16-60 nt in total, complementary stems, each in its own cubic box;
    temperatures spaced linearly over [t_lo, t_hi] Celsius."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        while True:
            s = int(rng.integers(6, 13))
            loop = int(rng.integers(4, 37))
            if 16 <= 2 * s + loop <= 60:
                break
        b = _Builder()
        add_hairpin(b, rng.integers(0, 4, s), rng.integers(0, 4, loop), (box / 2,) * 3, random_rotation(rng))
        tc = t_lo + (t_hi - t_lo) * i / max(1, n - 1)
        out.append(b.build([box] * 3, f"hairpin{i}", T=celsius(tc), salt=0.5))
    return out


def served_job(job_id: int, seed: int = 0, n_min: int = 40, n_max: int = 20000):
    """One job of the served stream: size log-uniform over [n_min, n_max]
    nucleotides, kind uniform over single strand / duplex / hairpin /
    duplex bundle, temperature uniform over 20-60 C, steps log-uniform
    over [1e4, 1e6].  Seeded per job."""
    rng = np.random.default_rng([seed, job_id])
    N = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
    kind = ["single", "duplex", "hairpin", "bundle"][int(rng.integers(0, 4))]
    tc = float(rng.uniform(20.0, 60.0))
    steps = int(round(math.exp(rng.uniform(math.log(1e4), math.log(1e6)))))
    b = _Builder()
    if kind == "single" or N < 16:
        n = max(N, 2)
        side = math.ceil(math.ceil(n / 60) ** (1.0 / 3.0))
        L = 30.0 * side
        _chunks_single(b, rng, n, L)
    elif kind == "hairpin" and N <= 200:
        s = max(3, min(N // 4, 12))
        loop = max(4, N - 2 * s)
        add_hairpin(b, rng.integers(0, 4, s), rng.integers(0, 4, loop), (0.0, 0.0, 0.0), random_rotation(rng))
        L = max(30.0, 2.0 * (loop * FENE_R0 / math.pi + 4.0))
        p = np.array(b.pos)
        b.pos = [x + L / 2 for x in p]
    else:
        # duplexes of at most 96 bp, bundled on a lattice when there are many
        bp_total = max(N // 2, 4)
        n_dup = max(1, math.ceil(bp_total / 96))
        bp = max(4, bp_total // n_dup)
        side = max(1, math.ceil(math.sqrt(n_dup)))
        spacing = 2.6 if kind == "bundle" else 6.0
        dy = spacing * math.sqrt(3.0) / 2.0
        Lz = bp * BASE_BASE + 4.0
        Lx, Ly = max(side * spacing, 12.0), max(side * dy, 12.0)
        k = 0
        for j in range(side):
            for i in range(side):
                if k == n_dup:
                    break
                c = np.array([(i + 0.5 * (j % 2) + 0.25) * spacing, (j + 0.5) * dy, Lz / 2])
                add_duplex(b, rng.integers(0, 4, bp), c, None)
                k += 1
        L = None
        sys_ = b.build([Lx, Ly, Lz], f"job{job_id}", T=celsius(tc), salt=0.5)
        return sys_, steps
    sys_ = b.build([L, L, L], f"job{job_id}", T=celsius(tc), salt=0.5)
    return sys_, steps


def _chunks_single(b: _Builder, rng, n, L):
    """Single strands of at most 60 nt scattered on a lattice, one strand
    per lattice cell; the n nucleotides are spread as evenly as possible
    (the first n mod n_str strands take one more), so that no strand is
    longer than its cell."""
    n_str = max(1, math.ceil(n / 60))
    per, extra = divmod(n, n_str)
    side = math.ceil(n_str ** (1.0 / 3.0))
    a = L / side
    k = 0
    for ix in range(side):
        for iy in range(side):
            for iz in range(side):
                if k == n_str:
                    return
                m = per + (1 if k < extra else 0)
                add_single_strand(b, rng.integers(0, 4, m), (np.array([ix, iy, iz]) + 0.5) * a, random_rotation(rng))
                k += 1


def backbone_sites(s: System) -> np.ndarray:
    return s.pos + BACK1 * s.a1 + BACK2 * s.a2
