"""The C ABI boundary and the host-side file formats, on the CPU: the HIP
library loads and exports every entry point include/oxh.h declares, the
ctypes mirror of oxh_system_desc has the C layout, topology / configuration
readers and writers agree with the reference's formats, the synthetic
builders produce valid oxDNA geometry, bench.py refuses several GPUs."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from _common import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "oxh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(oxh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from oxdna_amd import _lib
    assert os.path.exists(_lib.LIB_PATH), "liboxh.so is not built"
    L = C.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_lib.EXPORTS)


@pytest.mark.parametrize("py_name, c_name", [("SystemDesc", "oxh_system_desc"), ("ExtForce", "oxh_ext_force")])
def test_struct_layouts_match_c(tmp_path, py_name, c_name):
    from oxdna_amd import _lib
    cls = getattr(_lib, py_name)
    fields = [f[0] for f in cls._fields_]
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{ROOT}/include/oxh.h"', "int main(void) {",
             f'  printf("%zu\\n", sizeof({c_name}));']
    for f in fields:
        lines.append(f'  printf("%zu\\n", offsetof({c_name}, {f}));')
    lines.append("  return 0; }")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == C.sizeof(cls)
    for f, off in zip(fields, out[1:]):
        assert getattr(cls, f).offset == off, f


def test_topology_formats_round_trip(tmp_path):
    from oxdna_amd import io, systems
    s = systems.hairpin_set(2, n=3)[1]
    p = tmp_path / "old.top"
    io.write_topology(str(p), s, new_style=False)
    bt, n3, n5, st = io.read_topology(str(p))
    assert np.array_equal(bt, s.btype) and np.array_equal(n3, s.n3) and np.array_equal(n5, s.n5)
    # new style lists every strand 5'->3' and numbers particles in that
    # order; these builders number strands 3'->5', so a strand comes back
    # reversed, with n3 / n5 swapped accordingly
    p = tmp_path / "new.top"
    io.write_topology(str(p), s, new_style=True)
    bt, n3, n5, st = io.read_topology(str(p))
    assert np.array_equal(bt, s.btype[::-1])
    assert n5[0] == -1 and n3[-1] == -1 and np.all(n3[:-1] == np.arange(1, s.N))


def test_reference_topology_and_conf_files():
    from oxdna_amd import io
    bt, n3, n5, st = io.read_topology(os.path.join(GOLDEN, "ff_init.top"))
    assert len(bt) == 16 and st.max() == 2
    # "TGCA": T G C A listed 5'->3'; the 5' end has no n5
    assert list(bt[:4]) == [3, 1, 2, 0] and n5[0] == -1 and n3[0] == 1 and n3[3] == -1
    step, box, pos, a1, a2, a3, vel, L = io.read_conf(os.path.join(GOLDEN, "ff_init.dat"), 16)
    assert step == 10000 and np.allclose(box, 20)
    assert np.allclose(np.einsum("ij,ij->i", a1, a3), 0, atol=1e-12)
    assert np.allclose(np.linalg.norm(a2, axis=1), 1)


def test_conf_round_trip(tmp_path):
    from oxdna_amd import io, systems
    s = systems.duplex_20bp(1)
    s.vel = np.random.default_rng(0).normal(size=(s.N, 3))
    s.L = np.random.default_rng(1).normal(size=(s.N, 3))
    io.write_topology(str(tmp_path / "t.top"), s)
    io.write_conf(str(tmp_path / "c.dat"), s, step=123)
    t = io.load_system(str(tmp_path / "t.top"), str(tmp_path / "c.dat"))
    assert np.array_equal(t.pos, s.pos) and np.allclose(t.a1, s.a1, atol=1e-15) and np.array_equal(t.vel, s.vel)


def test_input_file_parser():
    from oxdna_amd import io
    d = io.parse_input("a = 1\n# c\nT = 20C # x\nblock = {\n  name = x\n  k = { y = 1 }\n}\nb=2\n")
    assert d["a"] == "1" and d["T"] == "20C" and d["b"] == "2" and d["block"].startswith("{")
    assert abs(io.temperature("20C") - (293.15 * 0.1 / 300)) < 1e-15
    assert abs(io.temperature("300K") - 0.1) < 1e-15 and io.temperature("0.11") == 0.11


@pytest.mark.parametrize("builder", ["duplex", "hairpin", "bundle", "gas", "jobs"])
def test_synthetic_systems_have_valid_bonds(builder):
    from oxdna_amd import systems
    if builder == "duplex":
        ss = [systems.duplex_20bp(0)]
    elif builder == "hairpin":
        ss = systems.hairpin_set(0, n=16)
    elif builder == "bundle":
        ss = [systems.duplex_bundle(0, nx=4, ny=4, bp=24, lz=12.0)]
    elif builder == "gas":
        ss = [systems.duplex_gas(0, n_duplex=8, bp=24, density=1e-3)]
    else:
        ss = [systems.served_job(k, seed=3)[0] for k in range(12)]
    for s in ss:
        back = systems.backbone_sites(s)
        for i in range(s.N):
            if s.n3[i] >= 0:
                d = np.linalg.norm(back[s.n3[i]] - back[i])
                assert abs(d - 0.7564) < 0.25, (s.name, i, d)
        assert np.allclose(np.einsum("ij,ij->i", s.a1, s.a3), 0, atol=1e-10)
        assert set(np.unique(s.btype)) <= {0, 1, 2, 3}


def test_hairpin_sizes_follow_the_spec():
    from oxdna_amd import systems
    hs = systems.hairpin_set(7, n=256)
    sizes = [h.N for h in hs]
    assert min(sizes) >= 16 and max(sizes) <= 60
    T = [h.T for h in hs]
    assert abs(T[0] - systems.celsius(25)) < 1e-12 and abs(T[-1] - systems.celsius(75)) < 1e-12


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_served_pool_plan():
    """The served form's pools: a workgroup pool for small jobs, one grid
    pool for mid-size jobs, single-slot pools for the largest; within the
    HBM budget, and every job size from the served distribution has a
    class."""
    from oxdna_amd import serve
    pools, classes = serve.plan_pools(8 << 30, 20000)
    assert classes[0] == serve.SMALL_CAP and all(c == serve.SMALL_CAP for c in pools[0])
    assert max(pools[1]) == serve.MID_CAP and len(pools) == 4
    assert all(len(p) == 1 and p[0] == 20000 for p in pools[2:])
    total = sum(serve.footprint_bytes(sum(p), 128) for p in pools)
    assert total <= 8 << 30
    # a tight budget drops mid-size slots but keeps one
    pools2, _ = serve.plan_pools(serve.footprint_bytes(serve.SMALL_CAP * 256) + 3 * serve.footprint_bytes(20000, 128), 20000)
    assert 1 <= len(pools2[1]) < len(pools[1])
    # nothing above MID_CAP: no large pools
    pools3, classes3 = serve.plan_pools(8 << 30, 4000)
    assert len(pools3) == 2 and classes3 == [serve.SMALL_CAP, serve.MID_CAP]


def test_quat_from_frame_reproduces_the_axes():
    """The quaternion layout oxh_interaction_compute_forces reads: rotating
    the unit vectors by q gives back a1, a2 = a3 x a1 and a3 (the kernels'
    own axis formulas), for frames that exercise every branch of the
    conversion (trace and each diagonal element largest)."""
    from oxdna_amd import backend
    rng = np.random.default_rng(3)
    # random rotations plus the four branch cases: identity and 180-degree turns
    q0 = rng.normal(size=(200, 4))
    q0 /= np.linalg.norm(q0, axis=1, keepdims=True)
    special = np.array([[0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0.]])
    q0 = np.vstack([q0, special])

    def axes(q):
        x, y, z, w = q.T
        a1 = np.stack([x * x - y * y - z * z + w * w, 2 * (x * y + z * w), 2 * (x * z - y * w)], axis=1)
        a3 = np.stack([2 * (x * z + y * w), 2 * (y * z - x * w), -x * x - y * y + z * z + w * w], axis=1)
        return a1, a3

    a1, a3 = axes(q0)
    q = backend.quat_from_frame(a1, a3)
    b1, b3 = axes(q)
    assert np.abs(np.linalg.norm(q, axis=1) - 1).max() < 1e-14
    assert np.abs(b1 - a1).max() < 1e-14 and np.abs(b3 - a3).max() < 1e-14
    # q and -q are the same rotation
    assert np.all(np.minimum(np.abs(q - q0).max(axis=1), np.abs(q + q0).max(axis=1)) < 1e-14)
