"""Build ``libspp_hip.so`` (the HIP C-ABI library) in-tree for gfx950.

Usage: ``python -m sigpyproc3_amd.build`` (or ``build()`` from Python). Each ``.hip``
source under ``csrc/`` is compiled to an object with hipcc, then linked into
``sigpyproc3_amd/libspp_hip.so``. hipcc cross-compiles gfx950 without a GPU present.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libspp_hip.so"
ARCH = os.environ.get("SPP_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for all device code (the compiler forms them
# from ordinary float code on gfx950); the host pass ignores the feature.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

# Per-source additions. The sweep kernel's inner loop is bound by LDS reads; the ILP
# scheduling strategy keeps a dozen of them in flight per wave (the default keeps 2-4) within
# the same register budget.
SOURCE_FLAGS = {
    "sweep.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"],
    "batch.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"],
}


def _hipcc() -> str:
    cand = os.environ.get("HIPCC") or shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return cand


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [_hipcc(), *HIP_FLAGS, *SOURCE_FLAGS.get(src.name, []), "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    """Compile every HIP source and link the shared library; returns its path."""
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
