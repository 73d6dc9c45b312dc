"""Builds libpystarc_hip.so (all HIP sources in csrc/) for gfx950, in-tree.

Usage: ``python -m pystarc_amd.build`` or ``pystarc_amd.build.build()``.
hipcc cross-compiles without a GPU; the shared library lands next to this
file so it travels with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libpystarc_hip.so"
ARCH = os.environ.get("PST_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-variable", "-Wno-unused-value",
]


def sources():
    return sorted(CSRC.glob("*.hip"))


def build(verbose: bool = True, jobs: int = 8) -> Path:
    """Compile every csrc/*.hip to an object (in parallel) and link the .so."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    t0 = time.time()
    procs = []
    objs = []
    for src in sources():
        obj = objdir / (src.stem + ".o")
        objs.append(obj)
        cmd = [hipcc, *FLAGS, "-c", str(src), "-o", str(obj)]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        while sum(p.poll() is None for _, p in procs) >= jobs:
            time.sleep(0.05)
    failed = []
    for src, p in procs:
        out = p.communicate()[0].decode(errors="replace")
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()]
            if lines:
                print(f"[build] {src.name}:\n" + "\n".join(lines[-20:]))
    if failed:
        msg = "\n".join(f"--- {s.name}\n{o}" for s, o in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"[build] {LIB.name} built in {time.time() - t0:.1f}s")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
