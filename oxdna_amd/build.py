"""Builds the gfx950 HIP library oxdna_amd/liboxh.so with hipcc.

    python -m oxdna_amd.build          # or: python oxdna_amd/build.py

One shared object from every oxdna_amd/csrc/*.hip, compiled for gfx950
only, with packed-FP32 instruction selection disabled
(-target-feature -packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "liboxh.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# compile flags for every source; packed FP32 instructions are switched off
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast", "-Wno-unused-result",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(verbose: bool = True) -> str:
    """Compiles every source into liboxh.so (always rebuilds)."""
    objs = []
    procs = []
    for src in sources():
        obj = os.path.join(CSRC, os.path.basename(src)[:-4] + ".o")
        cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = False
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"--- {os.path.basename(src)}\n{out}\n")
        elif verbose and out.strip():
            # the host pass warns that it does not know the packed-fp32 feature
            lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()]
            if lines:
                sys.stderr.write(f"--- {os.path.basename(src)}\n" + "\n".join(lines) + "\n")
    if failed:
        raise RuntimeError("hipcc failed")
    tmp = OUT + ".tmp"
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    for o in objs:
        os.remove(o)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build()
