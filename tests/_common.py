"""Helpers shared by the tests: golden fixtures, error metrics."""
from __future__ import annotations

import glob
import os

import numpy as np

from oxdna_amd import systems

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_names():
    return sorted(os.path.basename(f)[:-4] for f in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load_golden(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    s = systems.System(pos=z["pos"].copy(), a1=z["a1"].copy(), a3=z["a3"].copy(), btype=z["btype"].copy(),
                       n3=z["n3"].copy(), n5=z["n5"].copy(), strand=z["strand"].copy(), box=z["box"].copy(),
                       T=float(z["T"]), salt=float(z["salt"]), name=name)
    # the input-file options the reference ran with, under its key names
    opts = dict(dh_half_charged_ends=bool(z["half_charged_ends"]), dh_lambda=float(z["dh_lambda"]),
                dh_strength=float(z["dh_strength"]), debye_huckel_rhigh=float(z["dh_rhigh"]),
                hb_multiplier=float(z["hb_multiplier"]), max_backbone_force=float(z["max_backbone_force"]),
                max_backbone_force_far=float(z["max_backbone_force_far"]))
    if int(z["seqdep"]):
        opts.update(seq_stck=z["seq_stck"].copy(), seq_stck_fact_eps=float(z["seq_stck_fact_eps"]),
                    seq_hydr_AT=float(z["seq_hydr_AT"]), seq_hydr_GC=float(z["seq_hydr_GC"]))
    ref = dict(E8=z["E8"].copy(), F=z["F"].copy(), T=z["Tbody"].copy(), stress7=z["stress7"].copy(),
               hb_count=int(z["hb_count"]))
    return s, opts, ref


def oracle_params(s, opts):
    from oracle import pyoracle as po
    P = po.Params(T=s.T, salt=s.salt, half_charged_ends=opts.get("dh_half_charged_ends", True),
                  dh_lambda=opts.get("dh_lambda", -1.0), dh_strength=opts.get("dh_strength", -1.0),
                  dh_rhigh=opts.get("debye_huckel_rhigh", -1.0), hb_multiplier=opts.get("hb_multiplier", 1.0),
                  max_backbone_force=opts.get("max_backbone_force", -1.0),
                  max_backbone_force_far=opts.get("max_backbone_force_far", -1.0))
    if "seq_stck" in opts:
        P.set_seqdep(opts["seq_stck"], opts["seq_stck_fact_eps"], opts["seq_hydr_AT"], opts["seq_hydr_GC"])
    return P


def vec_rel_err(got, ref):
    """(max_i |d_i| / max_j |ref_j|, max_i |d_i| / max(|ref_i|, 0.1 * rms(ref)))
    for (N, 3) arrays: error relative to the largest vector, and
    per-particle error with a floor at a tenth of the RMS magnitude."""
    d = np.linalg.norm(got - ref, axis=1)
    m = np.linalg.norm(ref, axis=1)
    big = m.max() if m.size else 1.0
    floor = 0.1 * np.sqrt(np.mean(m ** 2)) if m.size else 1.0
    return float(d.max() / big), float(np.max(d / np.maximum(m, floor)))


R_CLOSE = 1.676   # close-pair bound of the lists (oxdna_amd/csrc/oxh_device.h)


def check_pair_set(got, t, P, skin=0.05, far_back=False, po=None):
    """A GPU pair set (unordered, i < j) against the FP64 Verlet predicate
    of the oracle (every pair with centres closer than r_cut + 2 skin).
    Without far_back the two are equal.  With far_back the list keeps the
    close pairs (centres within R_CLOSE + 2 skin) and those far pairs whose
    backbone sites are within the Debye-Hueckel cut-off plus 2 skin: pairs
    nearer than that bound must be present, pairs farther than it plus the
    filter's margin (1e-3 + 2^-18 of the longest box side, far above the
    FP32 error of wrapped sites) plus 1e-4 must be absent.
    Returns (predicate pairs, pairs kept)."""
    if po is None:
        from oracle import pyoracle as po
    rv = P.rcut + 2 * skin
    want = po.verlet_pairs(t.pos, t.n3, t.n5, t.box, rv * rv, max_pairs=max(1024, len(got) + 1024))
    if not far_back:
        assert got.shape == want.shape and np.array_equal(got, want)
        return len(want), len(got)
    a2 = np.cross(t.a3, t.a1)
    back = t.pos - 0.34 * t.a1 + 0.3408 * a2

    def dist(x, p):
        d = x[p[:, 1]] - x[p[:, 0]]
        d -= t.box * np.rint(d / t.box)
        return np.sqrt((d * d).sum(axis=1))

    wset = {tuple(x) for x in want}
    gset = {tuple(x) for x in got}
    assert gset <= wset, "pairs outside the Verlet predicate"
    dc, db = dist(t.pos, want), dist(back, want)
    bound = P.dh()["RC"] + 2 * skin
    must = (dc < R_CLOSE + 2 * skin - 1e-9) | (db < bound)
    margin = 1e-3 + np.ldexp(float(np.max(t.box)), -18) + 1e-4
    never = (dc >= R_CLOSE + 2 * skin + 1e-9) & (db >= bound + margin)
    keep = np.array([tuple(x) in gset for x in want], bool)
    assert np.all(keep[must]), f"{int((~keep[must]).sum())} pairs in range left out"
    assert not np.any(keep[never]), f"{int(keep[never].sum())} far pairs out of range kept"
    return len(want), len(got)
