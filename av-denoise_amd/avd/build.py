"""Builds ``libavd.so`` from ``av-denoise_amd/csrc`` with hipcc for gfx950, in-tree.

Each translation unit compiles in parallel into ``av-denoise_amd/_build``; the shared library lands
next to this file so it ships with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(PKG))
CSRC = os.path.join(os.path.dirname(PKG), "csrc")
BUILD = os.path.join(os.path.dirname(PKG), "_build")
LIB = os.path.join(PKG, "libavd.so")
ARCH = os.environ.get("AVD_OFFLOAD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", f"-I{os.path.join(ROOT, 'include')}",
    # Packed FP32 VALU ops are kept out of every kernel on this target.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# Exact-arithmetic units keep the operation order they spell out (explicit fmaf only).
EXACT = ["-ffp-contract=off"]
# Matrix-core accumulators in VGPRs rather than AGPRs: the grouping and filter kernels read every
# product's result with vector instructions, which AGPRs would first copy out.
MFMA_VGPR = ["-mllvm", "-amdgpu-mfma-vgpr-form"]
SOURCES = {
    "io.hip": EXACT,
    "motion.hip": EXACT,
    "collab.hip": EXACT + MFMA_VGPR,
    "nlm.hip": [],
    "noise.hip": EXACT,
    "grain.hip": EXACT,
    "grain_host.cpp": EXACT,
    "engine.cpp": EXACT,
    "host.cpp": EXACT,
    "tune.cpp": EXACT,
    "abi.cpp": EXACT,
}


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the HIP build needs ROCm")
    return path


def _compile(src: str, extra: list[str]) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{res.stderr}")
    return obj


def build(jobs: int | None = None) -> str:
    """Compiles every source and links ``libavd.so``; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as pool:
        objs = list(pool.map(lambda kv: _compile(*kv), SOURCES.items()))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
