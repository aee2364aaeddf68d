"""Statistical parity of the Bussi and Brownian ("john") thermostats on the
GPU, through the C ABI.

The reference pins its thermostats by equipartition (test/THERMOSTATS/
{BUSSI,JOHN}/quick_compare: <K> against 3T/2 per degree of freedom).  The
random streams here are Philox, not the reference's drand48 / curand, so
only distributions can agree.  The kinetic energy of f quadratic degrees of
freedom in the canonical ensemble is Gamma(f / 2, T) distributed:

  - Bussi (src/Backends/Thermostats/BussiThermostat.cpp:45-158) conserves
    the total momentum and rescales about the centre-of-mass velocity, so
    the translational kinetic energy relative to the centre of mass has
    f = 3(N - 1); the rotational one has f = 3N (unit moments of inertia).
    The reference's update advances its own stored kinetic energies, not
    the instantaneous ones, and then rescales the velocities to them: that
    stored value is a Markov chain whose stationary law is exactly
    Gamma(f / 2, T).  Sampled right after an activation, the system's
    kinetic energies equal the chain's values, so the Gamma law is tested
    exactly there (between activations the Newtonian steps trade kinetic
    for potential energy and only the mean is canonical);
  - Brownian (src/Backends/Thermostats/BrownianThermostat.cpp:43-68)
    redraws whole velocities and angular momenta from the Maxwell
    distribution (Andersen's scheme), which samples the canonical ensemble
    at any step: K_t (centre of mass included) and K_r have f = 3N.  Strong
    coupling (diff_coeff 0.01: p_t ~ 0.19 every 7 steps) so that 2,000
    steps equilibrate.

Samples: 128 independent 20-bp duplexes in one batch (each with its own
seed), 16 samples each, 500 steps apart.  Checked: the mean (equipartition)
within 4 standard errors, the variance within 4 of its standard errors,
and a Kolmogorov-Smirnov test against Gamma(f / 2, T).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THERMO = {
    "bussi": dict(thermostat="bussi", newtonian_steps=10, bussi_tau=100),
    "brownian": dict(thermostat="john", diff_coeff=0.01, newtonian_steps=7),
}
N_REPLICAS = 128
N_SAMPLES = 16


@pytest.fixture(scope="module")
def hip():
    from oxdna_amd import _lib, backend
    assert _lib.device_count() >= 1, "no GPU visible"
    return backend


def _samples(hip, monkeypatch, thermo, engine):
    from oxdna_amd import systems
    monkeypatch.setenv("OXH_PERSIST", "1" if engine == "workgroup" else "0")
    reps = [systems.duplex_20bp(3) for _ in range(N_REPLICAS)]
    T = reps[0].T
    per = [dict(seed=1000 + k) for k in range(N_REPLICAS)]
    kt, kr = [], []
    with hip.HIPBatch(reps, options=per, dt=0.003, verlet_skin=0.05, refresh_vel=True, **THERMO[thermo]) as b:
        assert b.stats()["engine"] == engine
        # Bussi activates at the end of steps 0, 10, 20, ... (step indices
        # from 0): after 2,001 + 500 k steps the last step was an activation
        b.run(2001)
        for _ in range(N_SAMPLES):
            b.run(500)
            for k in range(N_REPLICAS):
                st = b.state(k)
                v, L = st["vel"], st["L"]
                if thermo == "bussi":
                    v = v - v.mean(axis=0)
                kt.append(0.5 * (v ** 2).sum())
                kr.append(0.5 * (L ** 2).sum())
            assert b.info(0)[1] == 0
    n = reps[0].N
    ft = 3 * (n - 1) if thermo == "bussi" else 3 * n
    return np.array(kt), np.array(kr), ft, 3 * n, T


def _check_gamma(x, f, T, what):
    from scipy import stats
    k = f / 2.0
    m = len(x)
    mean_z = (x.mean() / (k * T) - 1.0) / np.sqrt(1.0 / (k * m))
    # relative standard error of a sample variance of a Gamma(k) variate:
    # sqrt((2 + 6 / k) / m)
    var_z = (x.var(ddof=1) / (k * T * T) - 1.0) / np.sqrt((2.0 + 6.0 / k) / m)
    ks = stats.kstest(x, stats.gamma(a=k, scale=T).cdf)
    print(f"{what}: f={f} <K>/(fT/2) {x.mean() / (k * T):.4f} (z {mean_z:+.2f}), "
          f"var ratio {x.var(ddof=1) / (k * T * T):.3f} (z {var_z:+.2f}), KS p {ks.pvalue:.3g}")
    assert abs(mean_z) < 4.0, (what, mean_z)
    assert abs(var_z) < 4.0, (what, var_z)
    assert ks.pvalue > 1e-3, (what, ks.pvalue)


@pytest.mark.parametrize("engine", ["workgroup", "grid"])
@pytest.mark.parametrize("thermo", ["bussi", "brownian"])
def test_kinetic_energy_is_canonical(hip, monkeypatch, thermo, engine):
    kt, kr, ft, fr, T = _samples(hip, monkeypatch, thermo, engine)
    _check_gamma(kt, ft, T, f"{thermo}/{engine} K_t")
    _check_gamma(kr, fr, T, f"{thermo}/{engine} K_r")
