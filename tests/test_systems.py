"""Synthetic workloads: the served stream's jobs start without overlaps.

A job whose strands overlap at t = 0 breaks a FENE bond within a few steps
on any backend (the reference's too), so a generator mistake shows up as a
failed job rather than as a slow one.  The check: no two backbone sites
that are not bonded neighbours sit closer than 0.6, inside the backbone
excluded-volume core (sigma = 0.7, src/model.h EXCL_S1).
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy.spatial import cKDTree

from oxdna_amd import systems


def closest_nonbonded_backbones(s) -> float:
    back = np.mod(s.pos + systems.BACK1 * s.a1 + systems.BACK2 * np.cross(s.a3, s.a1), s.box)
    pairs = cKDTree(back, boxsize=s.box).query_pairs(0.6, output_type="ndarray")
    if len(pairs) == 0:
        return np.inf
    keep = (s.n3[pairs[:, 0]] != pairs[:, 1]) & (s.n5[pairs[:, 0]] != pairs[:, 1])
    pairs = pairs[keep]
    if len(pairs) == 0:
        return np.inf
    d = back[pairs[:, 0]] - back[pairs[:, 1]]
    d -= s.box * np.rint(d / s.box)
    return float(np.sqrt((d * d).sum(axis=1)).min())


@pytest.mark.parametrize("seed", [1, 2])
def test_served_jobs_start_without_overlaps(seed):
    for k in range(64):
        s, _ = systems.served_job(k, seed=seed)
        assert closest_nonbonded_backbones(s) >= 0.6, f"seed {seed} job {k} (N {s.N}) starts with overlapping strands"


def test_single_strands_fit_their_cells():
    # 7,905 nt in strands of at most 60: 132 strands, none longer than 60
    s, _ = systems.served_job(2, seed=2)
    lengths = np.bincount(s.strand)
    assert s.N == 7905 and lengths.max() <= 60 and lengths.sum() == s.N
