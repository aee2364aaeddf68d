"""GPU tests of the host-driver features around the step, through the C ABI:
strand re-boxing (fix_diffusion), a broken bond reported as an error
instead of a crash, and the input-file driver run end to end against the
reference binary's own output files for the same input.
"""
import os
import shutil

import numpy as np
import pytest

from _common import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from oxdna_amd import _lib, backend
    assert _lib.device_count() >= 1, "no GPU visible"
    return backend


def test_fix_diffusion_reboxes_strands(hip):
    """Strands whose centre of mass lies outside the box come back by whole
    box lengths (src/Backends/SimBackend.cpp:788-882); the energy does not
    change (nearest images are invariant)."""
    from oxdna_amd import systems
    s = systems.duplex_gas(1, n_duplex=8, bp=12, density=1e-3)
    L = s.box
    t = s.copy()
    moves = {0: np.array([2, -1, 0]), 3: np.array([0, 0, 5]), 9: np.array([-3, 0, 1])}
    for strand, m in moves.items():
        t.pos[t.strand == strand] += m * L
    with hip.HIPBatch(t, seed=3) as b:
        E0 = b.energies()[0, 0]
        n = b.fix_diffusion()
        st = b.state()
        b.compute_forces()
        E1 = b.energies()[0, 0]
    assert n == len(moves)
    shift = (st["pos"] - t.pos) / L
    assert np.allclose(shift, np.rint(shift), atol=1e-9)
    for strand in np.unique(t.strand):
        sel = t.strand == strand
        com = st["pos"][sel].mean(axis=0)
        assert np.all(com >= 0) and np.all(com < L), (strand, com)
        # whole strands move together
        assert np.all(np.abs(shift[sel] - shift[sel][0]) < 1e-9)
    assert abs(E1 - E0) <= 1e-6 * abs(E0)


def test_broken_bond_is_reported(hip):
    """A bond stretched beyond the FENE range is an error of that system
    (the kernel's own check), not a crash."""
    from oxdna_amd import _lib, systems
    s = systems.duplex_20bp(0)
    s.pos[5] += np.array([0.0, 0.0, 1.5]) * s.a3[5] + 1.0 * s.a1[5]
    with hip.HIPBatch(s, thermostat="langevin", diff_coeff=2.5, seed=1) as b:
        with pytest.raises(_lib.OxhError, match="FENE"):
            b.run(10)
        nr, err = b.info(0)
        assert err & 1


@pytest.mark.parametrize("precision,e_tol,x_tol", [("double", 2e-6, 1e-10), ("mixed", 5e-6, 1e-5)])
def test_driver_follows_reference_nve(hip, tmp_path, precision, e_tol, x_tol):
    """The input file of tests/golden/driver_nve (300 NVE steps of a 20-bp
    duplex, thermal velocities) run by oxdna_amd.driver: the energy file
    (printed with 6 decimals), the trajectory and the last configuration
    follow the reference CPU binary's files for the same input.  Measured:
    FP64 forces give the same energy file and positions within 5e-14 after
    300 steps (bound 1e-10); FP32 forces 1e-6 and 2e-7 (bounds 5e-6,
    1e-5)."""
    from oxdna_amd import driver
    d = os.path.join(GOLDEN, "driver_nve")
    for f in ("input", "t.top", "c.dat"):
        shutil.copy(os.path.join(d, f), tmp_path / f)
    res = driver.run(str(tmp_path / "input"), [f"backend_precision = {precision}"])
    assert res["step"] == 300
    mine = np.loadtxt(tmp_path / "energy.dat", ndmin=2)
    ref = np.loadtxt(os.path.join(d, "energy_ref.dat"), ndmin=2)
    assert mine.shape == ref.shape
    de = np.abs(mine - ref).max()
    from test_driver import _confs
    traj, traj_ref = _confs(tmp_path / "trajectory.dat"), _confs(os.path.join(d, "trajectory_ref.dat"))
    last, last_ref = _confs(tmp_path / "last_conf.dat"), _confs(os.path.join(d, "last_conf_ref.dat"))
    assert [t for t, _ in traj] == [t for t, _ in traj_ref] and last[0][0] == last_ref[0][0] == 300
    dx = max(np.abs(a[:, :9] - b[:, :9]).max() for (_, a), (_, b) in zip(traj + last, traj_ref + last_ref))
    dv = max(np.abs(a[:, 9:] - b[:, 9:]).max() for (_, a), (_, b) in zip(traj + last, traj_ref + last_ref))
    print(f"{precision}: energy file {de:.1e}, positions/axes {dx:.1e}, velocities/L {dv:.1e}")
    assert de <= e_tol
    assert dx <= x_tol and dv <= 20 * x_tol
    # the data_output file (step, energy split, kinetic energy, HB count,
    # potential energy) against the reference's, column by column
    obs = np.loadtxt(tmp_path / "obs.dat", ndmin=2)
    obs_ref = np.loadtxt(os.path.join(d, "obs_ref.dat"), ndmin=2)
    assert obs.shape == obs_ref.shape and np.array_equal(obs[:, 0], obs_ref[:, 0])
    assert np.array_equal(obs[:, 10], obs_ref[:, 10])
    assert np.abs(obs[:, 1:10] - obs_ref[:, 1:10]).max() <= e_tol
    assert np.abs(obs[:, 11] - obs_ref[:, 11]).max() <= e_tol
    # the binary last configuration (lastconf_file_bin) against the
    # reference's, in its binary layout
    from oxdna_amd import io
    got = io.read_conf_binary(str(tmp_path / "last_conf.bin"), 40)
    want = io.read_conf_binary(os.path.join(d, "last_conf_ref.bin"), 40)
    assert got[0] == want[0] == 300 and np.array_equal(got[1], want[1])
    assert max(np.abs(a - b).max() for a, b in zip(got[2:6], want[2:6])) <= x_tol
    assert max(np.abs(a - b).max() for a, b in zip(got[6:8], want[6:8])) <= 20 * x_tol


def test_driver_reloads_reference_binary_configuration(hip, tmp_path):
    """reload_from with the reference's own reload file, a binary
    configuration (here its step-300 last configuration), under the
    reference's rules (no conf_file or seed, the step counter kept): the
    driver starts at step 300 with that state and prints the reference's
    step-300 energy line; 150 more NVE steps conserve the total energy."""
    from oxdna_amd import driver, io
    d = os.path.join(GOLDEN, "driver_nve")
    shutil.copy(os.path.join(d, "t.top"), tmp_path / "t.top")
    shutil.copy(os.path.join(d, "last_conf_ref.bin"), tmp_path / "reload.bin")
    inp = io.read_input(os.path.join(d, "input"))
    for key in ("conf_file", "seed", "lastconf_file_bin"):
        inp.pop(key, None)
    inp.update(reload_from="reload.bin", restart_step_counter="0", steps="150", backend_precision="double")
    with open(tmp_path / "input", "w") as f:
        for k, v in inp.items():
            f.write(f"{k} = {v}\n")
    res = driver.run(str(tmp_path / "input"))
    assert res["step"] == 450
    mine = np.loadtxt(tmp_path / "energy.dat", ndmin=2)
    ref = np.loadtxt(os.path.join(d, "energy_ref.dat"), ndmin=2)
    assert np.allclose(mine[0, 0], 0.9) and np.abs(mine[0] - ref[-1]).max() <= 2e-6
    assert np.abs(mine[:, 3] - mine[0, 3]).max() <= 1e-4
    with open(tmp_path / "input", "a") as f:
        f.write("seed = 3\n")
    with pytest.raises(ValueError, match="seed"):
        driver.run(str(tmp_path / "input"))


@pytest.mark.parametrize("case", ["grid_langevin_sorted", "grid_bussi", "workgroup", "fp64"])
def test_checkpoint_resumes_bitwise(hip, monkeypatch, case):
    """300 steps in one call equal 150 steps, a checkpoint, a new batch of
    the same systems restored from it and 150 more steps, bit for bit:
    positions, orientations, velocities, angular momenta, forces.  Covers
    Langevin noise (Philox counters), Bussi's internal kinetic energies and
    its step alignment, Hilbert re-sorts (the particle order is part of
    the state, sorts fall on the same steps), the workgroup engine and the
    FP64 form."""
    from oxdna_amd import systems
    opts = dict(dt=0.003, verlet_skin=0.05, seed=3)
    sort = 0
    if case == "workgroup":
        s = [h for h in systems.hairpin_set(5, n=8)][:4]
        opts.update(thermostat="langevin", diff_coeff=2.5)
    else:
        monkeypatch.setenv("OXH_PERSIST", "0")
        s = [systems.duplex_gas(5, n_duplex=64, bp=12, density=2e-3)]
        if case == "grid_bussi":
            opts.update(thermostat="bussi", newtonian_steps=10, bussi_tau=100)
        else:
            opts.update(thermostat="langevin", diff_coeff=2.5)
        if case == "grid_langevin_sorted":
            sort = 100
        if case == "fp64":
            opts.update(backend_precision="double")

    def final(b):
        out = []
        for k in range(len(s)):
            st = b.state(k)
            F, T, E = b.forces(k)
            out.append(dict(st, F=F, T=T, E=E))
        return out

    with hip.HIPBatch(s, **opts) as a:
        a.set_sort_interval(sort)
        a.run(300)
        want = final(a)
        engine = a.stats()["engine"]
    with hip.HIPBatch(s, **opts) as b:
        b.set_sort_interval(sort)
        b.run(150)
        ck = b.checkpoint()
    with hip.HIPBatch(s, **opts) as c:
        c.set_sort_interval(sort)
        c.restore(ck)
        c.run(150)
        got = final(c)
    assert engine == ("workgroup" if case == "workgroup" else "grid")
    for w, g in zip(want, got):
        for key in ("pos", "a1", "a3", "vel", "L", "F", "T", "E"):
            assert np.array_equal(w[key], g[key]), (case, key, np.abs(w[key] - g[key]).max())


def test_checkpoint_rejects_another_batch(hip):
    from oxdna_amd import _lib, systems
    with hip.HIPBatch(systems.duplex_20bp(0)) as a:
        ck = a.checkpoint()
    with hip.HIPBatch(systems.duplex_20bp(0), backend_precision="double") as b:
        with pytest.raises(_lib.OxhError, match="another batch|checkpoint"):
            b.restore(ck)


def test_driver_checkpoint_reload(hip, tmp_path):
    """The driver with checkpoint_every / checkpoint_file and reload_from
    (the reference's keys): 150 steps, a checkpoint, 150 more steps reloaded
    from it give the same last configuration, digit for digit, as 300
    steps in one run (Langevin thermostat, so the random streams resume
    too)."""
    from oxdna_amd import driver
    d = os.path.join(GOLDEN, "driver_nve")
    for f in ("input", "t.top", "c.dat"):
        shutil.copy(os.path.join(d, f), tmp_path / f)
    common = ["thermostat = langevin", "diff_coeff = 2.5", "print_conf_interval = 0", "trajectory_file = traj.dat"]
    driver.run(str(tmp_path / "input"), common + ["steps = 300", "lastconf_file = last_a.dat", "energy_file = e_a.dat"])
    driver.run(str(tmp_path / "input"), common + ["steps = 150", "lastconf_file = last_b1.dat", "energy_file = e_b1.dat",
                                                    "checkpoint_every = 150", "checkpoint_file = ck.bin"])
    res = driver.run(str(tmp_path / "input"), common + ["steps = 150", "lastconf_file = last_b2.dat", "energy_file = e_b2.dat",
                                                          "reload_from = ck.bin"])
    assert res["step"] == 300
    a = (tmp_path / "last_a.dat").read_text()
    b = (tmp_path / "last_b2.dat").read_text()
    assert a == b
    # the energy lines of the second half continue the first run's
    ea = (tmp_path / "e_a.dat").read_text().splitlines()
    eb = (tmp_path / "e_b1.dat").read_text().splitlines() + (tmp_path / "e_b2.dat").read_text().splitlines()[1:]
    assert ea == eb


@pytest.mark.parametrize("precision,e_tol,x_tol", [("double", 2e-6, 1e-10), ("mixed", 5e-6, 1e-5)])
def test_driver_external_forces_and_images(hip, tmp_path, precision, e_tol, x_tol):
    """tests/golden/driver_ext: 300 NVE steps from step 500 (the step
    counter is kept, so rates act) under every supported external force
    type, with the second strand starting one box length away, so that the
    run re-boxes it (fix_diffusion every 100 steps).  The energy file, the
    trajectory and the last configuration, printed with absolute positions
    (stored position + image count x box, as the reference prints them),
    follow the reference binary's; the binary last configuration carries
    the same positions and image counts."""
    from oxdna_amd import driver, io
    from test_driver import _confs
    d = os.path.join(GOLDEN, "driver_ext")
    for f in ("input", "t.top", "c.dat", "ext.txt"):
        shutil.copy(os.path.join(d, f), tmp_path / f)
    res = driver.run(str(tmp_path / "input"), [f"backend_precision = {precision}"])
    assert res["step"] == 800
    mine = np.loadtxt(tmp_path / "energy.dat", ndmin=2)
    ref = np.loadtxt(os.path.join(d, "energy_ref.dat"), ndmin=2)
    assert mine.shape == ref.shape
    de = np.abs(mine - ref).max()
    traj, traj_ref = _confs(tmp_path / "trajectory.dat"), _confs(os.path.join(d, "trajectory_ref.dat"))
    last, last_ref = _confs(tmp_path / "last_conf.dat"), _confs(os.path.join(d, "last_conf_ref.dat"))
    assert [t for t, _ in traj] == [t for t, _ in traj_ref]
    dx = max(np.abs(a[:, :9] - b[:, :9]).max() for (_, a), (_, b) in zip(traj + last, traj_ref + last_ref))
    print(f"{precision}: energy file {de:.1e}, positions/axes {dx:.1e}")
    assert de <= e_tol and dx <= x_tol
    got = io.read_conf_binary(str(tmp_path / "last_conf.bin"), 40, with_shift=True)
    want = io.read_conf_binary(os.path.join(d, "last_conf_ref.bin"), 40, with_shift=True)
    assert np.array_equal(got[-1], want[-1]) and want[-1].any()
    assert np.abs(got[2] - want[2]).max() <= x_tol
