"""GPU tests of slot pools (the served form) through the C ABI.

  - a job's trajectory is bitwise the same alone (a pool of one slot) and
    in a mix whose other jobs are admitted, run and released around it at
    other times (workgroup pools and grid pools with cell lists);
  - jobs stop exactly at their step targets while the others continue;
  - forces of a job in a pool equal the FP64 oracle's (FP32 tolerance);
  - a capacity change (a big job admitted into a pool whose lists must
    grow) does not change the running jobs' trajectories.
"""
import numpy as np
import pytest

from _common import vec_rel_err

pytestmark = pytest.mark.gpu

OPTS = dict(thermostat="langevin", diff_coeff=2.5, dt=0.003, verlet_skin=0.05, refresh_vel=True)


@pytest.fixture(scope="module")
def hip():
    from oxdna_amd import _lib, backend
    assert _lib.device_count() >= 1, "no GPU visible"
    return backend


def _jobs(lo, hi, count, seed=21):
    from oxdna_amd import systems
    out = []
    k = 0
    while len(out) < count:
        s, _ = systems.served_job(k, seed=seed)
        k += 1
        if lo <= s.N <= hi:
            out.append(s)
    return out


def _final(pool, slot):
    st = pool.state(slot)
    F, T, _ = pool.forces(slot)
    return dict(st, F=F, T=T)


def _alone(hip, cap, job, steps, seed, monkeypatch, lanes=None):
    if lanes:
        monkeypatch.setenv("OXH_FORCE_LANES", str(lanes))
    with hip.HIPPool([cap]) as p:
        p.admit(0, job, steps, seed=seed, **OPTS)
        p.run(steps + 100)          # the extra steps must not move it
        out = _final(p, 0)
        assert p.status()["done"][0] == steps
    return out


@pytest.mark.parametrize("kind", ["workgroup", "grid"])
def test_job_alone_equals_in_mix(hip, monkeypatch, kind):
    if kind == "workgroup":
        cap, jobs = 512, _jobs(40, 500, 6)
    else:
        cap, jobs = 4096, _jobs(1100, 4000, 4)   # > 1024 nt: cell lists
    steps = [700, 450, 900, 300, 650, 500][: len(jobs)]
    with hip.HIPPool([cap] * 3) as p:
        assert p.stats()["engine"] == kind
        # job 0 in slot 1 from the start; others come and go around it
        p.admit(1, jobs[0], steps[0], seed=500, **OPTS)
        p.admit(0, jobs[1], steps[1], seed=501, **OPTS)
        p.run(200)
        p.admit(2, jobs[2], steps[2], seed=502, **OPTS)
        p.run(137)
        st = p.status()
        assert st["done"][1] == 337 and st["done"][0] == 337 and st["done"][2] == 137
        p.run(200)   # job 1 reaches its target (450) inside this call
        st = p.status()
        assert st["done"][0] == steps[1] and st["done"][1] == 537
        p.release(0)
        p.admit(0, jobs[3], steps[3], seed=503, **OPTS)
        p.run(1000)
        st = p.status()
        assert list(st["done"]) == [steps[3], steps[0], steps[2]], st
        mixed0 = _final(p, 1)
        mixed2 = _final(p, 2)
    alone0 = _alone(hip, cap, jobs[0], steps[0], 500, monkeypatch)
    alone2 = _alone(hip, cap, jobs[2], steps[2], 502, monkeypatch)
    for a, m in ((alone0, mixed0), (alone2, mixed2)):
        for key in ("pos", "a1", "a3", "vel", "L", "F", "T"):
            assert np.array_equal(a[key], m[key]), (kind, key, np.abs(a[key] - m[key]).max())


@pytest.mark.parametrize("kind", ["workgroup", "grid"])
def test_pool_forces_match_oracle(hip, kind):
    from oracle import pyoracle as po
    cap, jobs = (512, _jobs(40, 500, 3)) if kind == "workgroup" else (4096, _jobs(1100, 4000, 2))
    with hip.HIPPool([cap] * len(jobs)) as p:
        for k, j in enumerate(jobs):
            p.admit(k, j, 300, seed=7 + k, **OPTS)
        p.run(300)
        for k, j in enumerate(jobs):
            st = p.state(k)
            F, T, E = p.forces(k)
            s = j.copy()
            s.pos, s.a1, s.a3 = st["pos"], st["a1"], st["a3"]
            ref = po.forces(po.Params(T=j.T, salt=j.salt), s)
            assert vec_rel_err(F, ref["F"])[0] <= 1e-5
            assert vec_rel_err(T, ref["T"])[0] <= 1e-5
            scale = np.abs(ref["E8"]).sum() + 1e-6 * s.N
            assert abs(E.sum() - ref["U"]) <= 1e-5 * scale


def test_capacity_growth_keeps_trajectories(hip, monkeypatch):
    """A dense bundle admitted next to a running duplex forces the
    neighbour rows to grow; the duplex's trajectory must not notice."""
    from oxdna_amd import systems
    small = _jobs(1100, 2000, 1, seed=5)[0]
    dense = systems.duplex_bundle(2, nx=6, ny=6, bp=48, lz=24.0)     # 3,456 nt at nucleotide density ~0.9
    with hip.HIPPool([4096, 4096]) as p:
        p.admit(0, small, 600, seed=11, **OPTS)
        p.run(250)
        p.admit(1, dense, 100, seed=12, **OPTS)
        p.run(500)
        mixed = _final(p, 0)
        assert p.status()["done"][0] == 600
    alone = _alone(hip, 4096, small, 600, 11, monkeypatch)
    for key in ("pos", "a1", "a3", "vel", "L", "F", "T"):
        assert np.array_equal(alone[key], mixed[key]), key
