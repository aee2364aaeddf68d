"""GPU tests of the two MD engines behind oxh_batch_run, through the C ABI:
the one-kernel-per-phase grid path and the workgroup-resident path (one
workgroup per system, many steps per launch, used when every system of the
batch has at most 512 nucleotides and no cells).

  - with the same lanes per nucleotide both engines give the same forces
    (bitwise) and the same trajectories up to FP32 rounding of the torque
    sums, for every thermostat;
  - a system's trajectory is the same alone and inside a batch of
    different systems (workgroup engine: lanes depend on the system only);
  - forces the engines leave behind after a run equal the FP64 oracle's on
    the downloaded state (FP32 tolerance of the force tests);
  - Langevin keeps <K_t> and <K_r> at 3NT/2 over a long run of small
    systems (equipartition, statistical; ~4 sigma band of the batch mean).
"""
import numpy as np
import pytest

from _common import vec_rel_err

pytestmark = pytest.mark.gpu

THERMO = {
    "langevin": dict(thermostat="langevin", diff_coeff=2.5),
    "bussi": dict(thermostat="bussi", newtonian_steps=10, bussi_tau=100),
    "brownian": dict(thermostat="john", diff_coeff=2.5, newtonian_steps=7),
    "nve": dict(thermostat="no"),
}


@pytest.fixture(scope="module")
def hip():
    from oxdna_amd import _lib, backend
    assert _lib.device_count() >= 1, "no GPU visible"
    return backend


def _hairpins(lo, hi, count, seed=11):
    from oxdna_amd import systems
    out = [h for h in systems.hairpin_set(seed, n=4 * count) if lo <= h.N <= hi]
    assert len(out) >= count
    return out[:count]


def _run(hip, monkeypatch, hs, steps, engine, lanes=None, **opts):
    monkeypatch.setenv("OXH_PERSIST", "1" if engine == "workgroup" else "0")
    if lanes is not None:
        monkeypatch.setenv("OXH_FORCE_LANES", str(lanes))
        monkeypatch.setenv("OXH_FAR_LANES", "0")
    per = [dict(seed=100 + k) for k in range(len(hs))]
    with hip.HIPBatch(hs, options=per, dt=0.003, verlet_skin=0.05, refresh_vel=True, **opts) as b:
        assert b.stats()["engine"] == engine
        b.run(steps)
        out = []
        for k in range(len(hs)):
            st = b.state(k)
            F, T, E = b.forces(k)
            out.append(dict(st, F=F, T=T, E=E, rebuilds=b.info(k)[0]))
    monkeypatch.delenv("OXH_FORCE_LANES", raising=False)
    monkeypatch.delenv("OXH_FAR_LANES", raising=False)
    return out


@pytest.mark.parametrize("thermo", list(THERMO))
def test_engines_agree(hip, monkeypatch, thermo):
    """Same lanes per nucleotide (8 for 33-64 nt): forces of the same state
    are bitwise equal and torques equal to FP32 rounding (the compiler may
    fuse multiply-adds of the torque sums differently in the two kernels).
    20 steps later the trajectories agree to 1e-6: the torque rounding
    moves L by ~dt * ulp(|T|) ~ 1e-9 per step; a wrong term moves it by
    ~1e-2."""
    hs = _hairpins(33, 64, 12)
    a0 = _run(hip, monkeypatch, hs, 0, "grid", lanes=8, **THERMO[thermo])
    b0 = _run(hip, monkeypatch, hs, 0, "workgroup", **THERMO[thermo])
    for x, y in zip(a0, b0):
        assert np.array_equal(x["F"], y["F"]) and np.array_equal(x["E"], y["E"])
        assert np.abs(x["T"] - y["T"]).max() <= 1e-6 * np.abs(x["T"]).max()
    a = _run(hip, monkeypatch, hs, 20, "grid", lanes=8, **THERMO[thermo])
    b = _run(hip, monkeypatch, hs, 20, "workgroup", **THERMO[thermo])
    for k, (x, y) in enumerate(zip(a, b)):
        assert x["rebuilds"] == y["rebuilds"], (k, x["rebuilds"], y["rebuilds"])
        for key in ("pos", "a1", "a3", "vel", "L"):
            assert np.abs(x[key] - y[key]).max() <= 1e-6, (thermo, k, key, np.abs(x[key] - y[key]).max())


def test_alone_equals_in_batch(hip, monkeypatch):
    from oxdna_amd import systems
    hs = systems.hairpin_set(5, n=24)
    mixed = _run(hip, monkeypatch, hs, 400, "workgroup", **THERMO["langevin"])
    for k in (0, 7, 23):
        monkeypatch.setenv("OXH_PERSIST", "1")
        with hip.HIPBatch(hs[k], dt=0.003, verlet_skin=0.05, refresh_vel=True, seed=100 + k, **THERMO["langevin"]) as b:
            b.run(400)
            st = b.state(0)
        for key in ("pos", "a1", "a3", "vel", "L"):
            assert np.array_equal(st[key], mixed[k][key]), (k, key)


@pytest.mark.parametrize("engine", ["grid", "workgroup"])
def test_forces_after_run_match_oracle(hip, monkeypatch, engine):
    from oracle import pyoracle as po
    hs = _hairpins(16, 60, 6, seed=3)
    res = _run(hip, monkeypatch, hs, 500, engine, **THERMO["langevin"])
    for h, r in zip(hs, res):
        s = h.copy()
        s.pos, s.a1, s.a3 = r["pos"], r["a1"], r["a3"]
        ref = po.forces(po.Params(T=h.T, salt=h.salt), s)
        assert vec_rel_err(r["F"], ref["F"])[0] <= 1e-5
        assert vec_rel_err(r["T"], ref["T"])[0] <= 1e-5
        scale = np.abs(ref["E8"]).sum() + 1e-6 * s.N
        assert abs(r["E"].sum() - ref["U"]) <= 1e-5 * scale


def test_langevin_equipartition_small_systems(hip, monkeypatch):
    """<K_t> = <K_r> = 3NT/2 under Langevin (diff_coeff 0.1: gamma_t = T/D
    ~ 1, relaxation in ~330 steps) for 64 hairpins at 30 C; reference
    thermostat: src/Backends/Thermostats/LangevinThermostat.cpp:57-85."""
    from oxdna_amd import systems
    hs = systems.hairpin_set(9, n=64, t_lo=30.0, t_hi=30.0)
    monkeypatch.setenv("OXH_PERSIST", "1")
    per = [dict(seed=7 + k) for k in range(len(hs))]
    n = sum(h.N for h in hs)
    kt, kr = [], []
    with hip.HIPBatch(hs, options=per, dt=0.003, verlet_skin=0.05, refresh_vel=True, thermostat="langevin",
                      diff_coeff=0.1) as b:
        b.run(3000)
        for _ in range(40):
            b.run(100)
            E = b.energies()
            kt.append(E[:, 2].sum() / (1.5 * n * hs[0].T))
            kr.append(E[:, 3].sum() / (1.5 * n * hs[0].T))
    # one sample of the batch has sigma ~ sqrt(2 / 3N) ~ 1.7 %; 40 samples,
    # correlated over ~3 samples: sigma of the mean ~ 0.5 %
    assert abs(np.mean(kt) - 1.0) < 0.02, np.mean(kt)
    assert abs(np.mean(kr) - 1.0) < 0.03, np.mean(kr)
