"""The reference program itself running on the GPU through the HIPDNA2
interaction plugin (oxdna_amd/plugin/HIPDNA2.cpp): the reference binary
built from its own sources (oracle/_ref/oxDNA, strict FP64 build) runs the
input files of tests/golden/driver_nve and driver_ext with
`interaction_type = HIPDNA2` and `plugin_search_path` naming the plugin,
and its output files are compared with the files the same binary wrote for
`interaction_type = DNA2_nomesh` on the CPU.

What runs where: the reference's MD_CPUBackend integrates, applies the
external forces (driver_ext: string, trap and mutual_trap forces, set on
each particle before the pair loop) and re-boxes strands (fix_diffusion,
whose energy checks call get_system_energy); every pair force, torque and
energy comes from liboxh's kernels, added onto the external forces.

Bars (the driver tests' bars): energy file (6 printed decimals) within
2e-6 in FP64 and 5e-6 in FP32; positions and axes of the trajectory and
last configuration within 1e-10 (FP64) and 1e-5 (FP32).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from _common import GOLDEN

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLUGIN_DIR = os.path.join(ROOT, "oxdna_amd", "plugin")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "oxDNA")


@pytest.fixture(scope="module")
def ready():
    from oxdna_amd import _lib
    assert _lib.device_count() >= 1, "no GPU visible"
    if not os.path.exists(os.path.join(PLUGIN_DIR, "HIPDNA2.so")) or not os.path.exists(REF_BIN):
        pytest.skip("HIPDNA2.so or the reference binary was not built (built where the reference sources are)")


@pytest.mark.parametrize("case", ["driver_nve", "driver_ext"])
@pytest.mark.parametrize("precision,e_tol,x_tol", [("double", 2e-6, 1e-10), ("mixed", 5e-6, 1e-5)])
def test_reference_binary_runs_on_gpu_plugin(ready, tmp_path, case, precision, e_tol, x_tol):
    from test_driver import _confs
    d = os.path.join(GOLDEN, case)
    for f in os.listdir(d):
        if not f.endswith("_ref.dat") and not f.endswith("_ref.bin"):
            shutil.copy(os.path.join(d, f), tmp_path / f)
    text = (tmp_path / "input").read_text().replace("interaction_type = DNA2_nomesh", "interaction_type = HIPDNA2")
    text += f"\nplugin_search_path = {PLUGIN_DIR}\nbackend_precision = {precision}\n"
    (tmp_path / "input").write_text(text)
    print(f"running the reference binary with HIPDNA2 on {case} ({precision})", flush=True)
    r = subprocess.run([REF_BIN, "input"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    log = (tmp_path / "log.dat").read_text() if (tmp_path / "log.dat").exists() else ""
    assert r.returncode == 0, (r.stdout + r.stderr + log)[-3000:]
    mine = np.loadtxt(tmp_path / "energy.dat", ndmin=2)
    ref = np.loadtxt(os.path.join(d, "energy_ref.dat"), ndmin=2)
    assert mine.shape == ref.shape
    de = np.abs(mine - ref).max()
    last, last_ref = _confs(tmp_path / "last_conf.dat"), _confs(os.path.join(d, "last_conf_ref.dat"))
    traj, traj_ref = _confs(tmp_path / "trajectory.dat"), _confs(os.path.join(d, "trajectory_ref.dat"))
    assert [t for t, _ in traj] == [t for t, _ in traj_ref] and last[0][0] == last_ref[0][0]
    dx = max(np.abs(a[:, :9] - b[:, :9]).max() for (_, a), (_, b) in zip(traj + last, traj_ref + last_ref))
    dv = max(np.abs(a[:, 9:] - b[:, 9:]).max() for (_, a), (_, b) in zip(traj + last, traj_ref + last_ref))
    print(f"{case} {precision}: energy file {de:.1e}, positions/axes {dx:.1e}, velocities/L {dv:.1e}")
    assert de <= e_tol
    assert dx <= x_tol and dv <= 20 * x_tol


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_reference_binary_runs_rna_test_on_gpu_plugin(ready, tmp_path, precision):
    """The reference's own oxRNA2 test (test/RNA/FORCE_FIELD/AVG_SEQ: input
    file, topology, configuration; tests/golden/rna/ff16_input) run by the
    reference binary with `interaction_type = HIPRNA2`: the potential
    energy split it writes (its split_energy.dat, computed on the GPU)
    matches its published reference.dat (6 decimals; the HB column within
    the CPU mesh's departure from the analytic f4 the GPU evaluates)."""
    if not os.path.exists(os.path.join(PLUGIN_DIR, "HIPRNA2.so")):
        pytest.skip("HIPRNA2.so was not built (built where the reference sources are)")
    src = os.path.join(GOLDEN, "rna", "ff16_input")
    (tmp_path / "run").mkdir()
    for f in ("init.top", "init.dat"):
        shutil.copy(os.path.join(src, f), tmp_path / f)
    text = open(os.path.join(src, "run", "input")).read().replace("interaction_type = RNA2", "interaction_type = HIPRNA2")
    text += f"\nplugin_search_path = {PLUGIN_DIR}\nbackend_precision = {precision}\n"
    (tmp_path / "run" / "input").write_text(text)
    print(f"running the reference binary with HIPRNA2 ({precision})", flush=True)
    r = subprocess.run([REF_BIN, "input"], cwd=tmp_path / "run", capture_output=True, text=True, timeout=300)
    log = (tmp_path / "run" / "log.dat").read_text() if (tmp_path / "run" / "log.dat").exists() else ""
    assert r.returncode == 0, (r.stdout + r.stderr + log)[-3000:]
    got = np.loadtxt(tmp_path / "run" / "split_energy.dat", ndmin=2)[0]
    want = np.loadtxt(os.path.join(src, "run", "reference.dat"))
    print(precision, "got", got, "want", want)
    other = [0, 1, 2, 3, 5, 6, 7]
    assert np.all(np.abs(got[other] - want[other]) <= 1.01e-6)
    assert abs(got[4] - want[4]) <= 5e-5 * abs(want[4]) + 1.01e-6
