"""Compile libmlfma.so (host C++ + HIP kernels for gfx950) in-tree.

    python -m mlfma_pmchwt_amd.build [--incremental]

Objects go to mlfma_pmchwt_amd/_build/, the library to mlfma_pmchwt_amd/libmlfma.so.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SRC = PKG / "csrc"
OUT = PKG / "_build"
LIB = PKG / "libmlfma.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", f"-I{ROOT / 'include'}", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unknown-pragmas"]
# Device code: gfx950 only; packed-FP32 instructions disabled.
DEVICE = ["-x", "hip", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]
# Host-only translation units (hipcc would otherwise add a default-target device pass).
HOST = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _sources():
    return sorted(list(SRC.glob("*.cpp")) + list(SRC.glob("*.hip")))


def _headers():
    return sorted(list(SRC.glob("*.hpp")) + [ROOT / "include" / "mlfma.h"])


def _compile(src: Path, incremental: bool) -> Path:
    obj = OUT / (src.name + ".o")
    if incremental and obj.exists():
        newest = max([src.stat().st_mtime] + [h.stat().st_mtime for h in _headers()])
        if obj.stat().st_mtime >= newest:
            return obj
    cmd = [HIPCC] + COMMON + (DEVICE if src.suffix == ".hip" else HOST) + ["-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(incremental: bool = False, jobs: int | None = None) -> Path:
    OUT.mkdir(exist_ok=True)
    srcs = _sources()
    with ThreadPoolExecutor(max_workers=jobs or min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fopenmp", "--offload-arch=gfx950", "-o", str(tmp)] + [str(o) for o in objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cpp_caller()
    return LIB


def build_cpp_caller() -> Path:
    """The C++ drop-in caller test program (tests/cpp/dropin_gmres.cpp) against libmlfma.so."""
    exe = OUT / "dropin_gmres"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "cpp" / "dropin_gmres.cpp"),
           "-o", str(exe), f"-L{PKG}", "-lmlfma", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return exe


if __name__ == "__main__":
    print(build(incremental="--incremental" in sys.argv))
