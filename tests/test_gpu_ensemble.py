"""Ensemble parity with the reference binary: the 20-bp duplex (random ACGT
sequence, cubic box of 20, 0.5 M, dt 0.003, Langevin with diff_coeff 2.5)
sampled by the HIP path and by the reference's CPU backend under the same
protocol: independent runs of 3e5 steps from the same ideal helix with
fresh thermal velocities, a sample every 1000 steps, the first 5e4 steps
dropped.

The reference's series (tests/golden/ensemble/duplex20.npz, written by
tests/golden/make_ensemble_golden.py with the strict-IEEE reference build:
192 runs per temperature, 20 C and 50 C) give the means of the potential
energy per nucleotide and of the base-pair fraction (hydrogen-bonded pairs
with energy below -0.1, the reference's hb_list observable with only_count,
src/Observables/HBList.cpp:64-164, over 20 possible pairs).  The HIP path
runs 256 replicas of the same system in one batch, each with its own seed
(more runs than the reference's, so that the comparison's error is mostly
the reference's own).  The duplex's slow modes (fraying ends) decorrelate
over 1e5 steps or more, so each run contributes one mean and the standard
error comes from the spread of those independent run means.  The two codes'
means of U/N and of the base-pair fraction are compared through
z = (difference) / (combined standard error).  The four comparisons (two
quantities, two temperatures) are judged together at the 5 % level:
|z| <= 2.5 each (Bonferroni), where a 2-sigma bar on each would fail a
correct code one time in six; every z is printed.

The kinetic energy per nucleotide is checked against its closed form,
Langevin equipartition <K/N> = 3T (translational and rotational, 3T/2
each): the HIP path's within three standard errors.  The reference's own
runs deviate from 3T by +0.2 % (20 C) and -0.4 % (50 C), up to three of
their standard errors, so the fixture is only pinned to 1 % of 3T there.
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ensemble", "duplex20.npz")
BURN = 50      # samples (of 1000 steps) dropped at the start of each run
REPLICAS = 256
Z_MAX = 2.5    # four comparisons at a family-wise 5 % level


def run_stats(x):
    """Mean over independent runs of x (runs, samples) after the burn-in,
    and its standard error from the spread of the run means."""
    m = np.asarray(x, np.float64)[:, BURN:].mean(axis=1)
    return float(m.mean()), float(m.std(ddof=1) / np.sqrt(m.size))


def reference():
    z = np.load(GOLDEN)
    return z, [float(t) for t in z["temps_c"]]


def test_reference_ensemble_equipartition():
    from oxdna_amd import systems
    z, temps = reference()
    for ti, tc in enumerate(temps):
        T = systems.celsius(tc)
        m, se = run_stats(z["K"][ti])
        print(f"reference {tc:.0f} C: <K/N> {m:.5f} +- {se:.5f}, 3T = {3 * T:.5f}")
        assert abs(m - 3 * T) <= 0.01 * 3 * T, (tc, m, se, 3 * T)
        # the fixture is an equilibrium series of a bound duplex
        assert z["hb"][ti].min() >= 0 and z["hb"][ti].max() <= 20


@pytest.mark.gpu
@pytest.mark.parametrize("ti", [0, 1], ids=["20C", "50C"])
def test_ensemble_matches_reference(ti):
    from oxdna_amd import _lib, backend, systems
    assert _lib.device_count() >= 1, "no GPU visible"
    z, temps = reference()
    tc = temps[ti]
    s = systems.duplex_20bp(int(z["system_seed"]))
    s.T = systems.celsius(tc)
    every = int(z["every"])
    n = int(z["steps"]) // every + 1
    U = np.zeros((REPLICAS, n))
    K = np.zeros((REPLICAS, n))
    hb = np.zeros((REPLICAS, n))
    opts = [dict(seed=9000 + 37 * ti + r) for r in range(REPLICAS)]
    with backend.HIPBatch([s] * REPLICAS, options=opts, thermostat="langevin", diff_coeff=2.5, dt=0.003,
                          verlet_skin=0.05, refresh_vel=True, salt_concentration=0.5) as b:
        for k in range(n):
            if k:
                b.run(every)
            E = b.energies()
            U[:, k] = E[:, 0] / s.N
            K[:, k] = (E[:, 2] + E[:, 3]) / s.N
            hb[:, k] = b.hb_count() / 20.0
        nreb, err = b.info()
        assert err == 0
    mk, sk = run_stats(K)
    print(f"{tc:.0f} C K/N: HIP {mk:.5f} +- {sk:.5f}, 3T = {3 * s.T:.5f}, z = {(mk - 3 * s.T) / sk:+.2f}")
    assert abs(mk - 3 * s.T) <= 3 * sk, (mk, sk, 3 * s.T)
    for name, gpu, ref in (("U/N", U, z["U"][ti]), ("bp fraction", hb, z["hb"][ti] / 20.0)):
        mg, sg = run_stats(gpu)
        mr, sr = run_stats(ref)
        sig = np.hypot(sg, sr)
        print(f"{tc:.0f} C {name}: HIP {mg:.5f} +- {sg:.5f}, reference {mr:.5f} +- {sr:.5f}, "
              f"z = {(mg - mr) / sig:+.2f}")
        assert abs(mg - mr) <= Z_MAX * sig, (name, mg, sg, mr, sr)
