"""Build the HIP/C++ engine library in-tree: psierra_amd/libpsierra_amd.so.

hipcc cross-compiles for gfx950 without a GPU.  Every .hip source is built
with packed FP32 instructions disabled (``-Xclang -target-feature -Xclang
-packed-fp32-ops``).  Objects go to psierra_amd/_build/ (ignored by git).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libpsierra_amd.so"
ARCH = os.environ.get("PSX_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = sorted(p.name for p in CSRC.glob("*.hip"))
CPP_SOURCES = sorted(p.name for p in CSRC.glob("*.cpp"))

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", f"-I{PKG.parent / 'include'}"]
DEVICE = [
    "-x",
    "hip",
    f"--offload-arch={ARCH}",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-ffp-contract=fast",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the engine needs ROCm's hipcc to build")


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def build(verbose: bool = False) -> Path:
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        jobs.append([hipcc, *COMMON, *DEVICE, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    for src in CPP_SOURCES:
        obj = BUILD / (src + ".o")
        jobs.append([hipcc, *COMMON, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    tmp = LIB.with_suffix(".so.tmp")
    _run(
        [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread"],
        verbose,
    )
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
