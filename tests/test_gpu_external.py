"""External forces on the GPU (oxh_batch_set_external_forces) against the
reference binary: forces and torques of a state under every supported
force type (string, trap, mutual_trap, constant_trap, repulsion_plane,
sphere, lowdim_trap, attraction_plane, repulsion_plane_moving,
sphere_smooth, repulsive_sphere_moving, repulsive_kepler_poinsot, LJ_wall,
LJ_cone, twist, generic_central_force, ellipsoid, com, yukawa_sphere), with
particle lists and ranges, rates at a nonzero step counter, nearest images
and a moving plane's clamp (fixtures of tests/golden/make_external_golden.py:
ForceDump of the reference's CPU backend, which adds the external forces to
the interaction's).

Bars: the force tests' FP32 bar (1e-5 of the largest force, per particle
1e-5) in the mixed form, 1e-9 in the all-FP64 form; in the all-FP64 form
the external part alone (forces with minus forces without) within 1e-8 of
its largest value.  External forces carry no torque and no potential
energy in the energy split, so torques and the split equal those of the
interaction alone.
"""
import glob
import os

import numpy as np
import pytest

from _common import GOLDEN, vec_rel_err

pytestmark = pytest.mark.gpu

NAMES = sorted(os.path.basename(f)[:-4] for f in glob.glob(os.path.join(GOLDEN, "external", "*.npz")))


@pytest.fixture(scope="module")
def hip():
    from oxdna_amd import _lib, backend
    assert _lib.device_count() >= 1, "no GPU visible"
    return backend


def load(name, tmp_path):
    from oxdna_amd import io, systems
    z = np.load(os.path.join(GOLDEN, "external", name + ".npz"))
    s = systems.System(pos=z["pos"].copy(), a1=z["a1"].copy(), a3=z["a3"].copy(), btype=z["btype"].copy(),
                       n3=z["n3"].copy(), n5=z["n5"].copy(), strand=z["strand"].copy(), box=z["box"].copy(),
                       T=float(z["T"]), salt=float(z["salt"]), name=name)
    path = tmp_path / "ext.txt"
    path.write_text(str(z["forces_file"]))
    recs = io.read_external_forces(str(path), s)
    return s, recs, int(z["step"]), z


@pytest.mark.parametrize("precision", ["mixed", "double"])
@pytest.mark.parametrize("name", NAMES)
def test_external_forces_vs_reference(hip, tmp_path, name, precision):
    s, recs, step, z = load(name, tmp_path)
    with hip.HIPBatch(s, backend_precision=precision) as b:
        F0, T0, _ = b.forces()
        b.set_external_forces(recs, step0=step)
        F, T, Ep = b.forces()
        e8 = b.energy_split()
        assert b.stats()["engine"] != "workgroup"
    tol = 1e-5 if precision == "mixed" else 1e-9
    big, pp = vec_rel_err(F, z["F"])
    print(f"{name} {precision}: F {big:.2e}/{pp:.2e}")
    assert big <= tol and pp <= tol
    assert vec_rel_err(T, z["Tbody"])[0] <= tol
    # no torque from the external forces: the interaction's torques (a small
    # mixed-form batch leaves the workgroup engine when forces are set, and
    # the grid engine's lanes per nucleotide sum in another order)
    if precision == "double":
        assert np.array_equal(T, T0)
    else:
        assert vec_rel_err(T, T0)[0] <= 1e-6
    scale = np.abs(z["E8"]).sum()
    assert np.all(np.abs(e8 - z["E8"]) <= tol * scale)
    # without the external forces the forces are the interaction's alone
    assert vec_rel_err(F0, z["F_pairs"])[0] <= tol
    if precision == "double":
        ext, ext_ref = F - F0, z["F"] - z["F_pairs"]
        err = np.abs(ext - ext_ref).max() / np.abs(ext_ref).max()
        print(f"{name}: external part {err:.2e}")
        assert err <= 1e-8


def test_external_forces_follow_steps_and_sorts(hip, tmp_path):
    """A run with external forces: the step counter advances with the
    steps (forces after 200 steps from step 800 equal those of the same
    state with the forces set anew at step 1000), Hilbert re-sorts keep each record on its
    particle, and removing the forces restores the interaction's."""
    s, recs, step, z = load("ext_duplex_step1000", tmp_path)
    with hip.HIPBatch(s, thermostat="langevin", diff_coeff=2.5, seed=3) as b:
        b.set_external_forces(recs, step0=800)
        b.set_sort_interval(64)
        b.run(200)
        F_run, _, _ = b.forces()
        st = b.state()
        # the last force evaluation read 800 + 199; set anew with the
        # counter of the next step (1000), the forces of the current state
        # are evaluated at 999 again
        b.set_external_forces(recs, step0=1000)
        F_set, _, _ = b.forces()
        b.set_external_forces([], step0=0)
        F_none, _, _ = b.forces()
    assert np.abs(F_run - F_set).max() <= 1e-6 * np.abs(F_set).max()
    t = s.copy()
    t.pos, t.a1, t.a3 = st["pos"], st["a1"], st["a3"]
    with hip.HIPBatch(t) as c:
        F_pairs, _, _ = c.forces()
    assert np.abs(F_none - F_pairs).max() <= 1e-6 * np.abs(F_pairs).max()
