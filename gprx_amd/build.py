"""Builds libgprx_hip.so (gfx950) in-tree from gprx_amd/csrc/*.hip.

Each translation unit is compiled by hipcc in parallel, then linked into one
shared library next to this file.  Run as a script or call build().
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OBJ = HERE / "_build"
LIB = HERE / "libgprx_hip.so"
RNG_SRC = HERE / "host" / "seeded_rng.cpp"
RNG_LIB = HERE / "libgprx_rng.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "-std=c++17",
    "-O3",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-Wno-unused-command-line-argument",
    # Keep the compiler from forming packed f32 arithmetic on its own.
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ".o")
    deps = [src] + sorted(CSRC.glob("*.hpp")) + [HERE.parent / "include" / "gprx_hip.h"]
    if obj.exists() and all(obj.stat().st_mtime >= d.stat().st_mtime for d in deps):
        return obj
    cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        msg = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"hipcc failed on {src.name}:\n{msg}")
    return obj


def build(force: bool = False, jobs: int | None = None) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    if force:
        for o in OBJ.glob("*.o"):
            o.unlink()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    build_rng()
    return LIB


def build_rng() -> Path:
    """Host-only SeededRng used to draw the synthetic workloads (g++)."""
    tmp = RNG_LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-o", str(tmp), str(RNG_SRC)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed on {RNG_SRC.name}:\n{res.stderr}")
    os.replace(tmp, RNG_LIB)
    return RNG_LIB


if __name__ == "__main__":
    path = build(force="--force" in sys.argv)
    print(path)
