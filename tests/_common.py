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
