"""Build libsv.so (HIP kernels for gfx950 + the C ABI) in-tree.

Usage: python skyfall-viewer_amd/build.py [--jobs N]

Every HIP translation unit is compiled with
  -ffp-contract=off                        keep the written float operation order
  -fhip-fp32-correctly-rounded-divide-sqrt IEEE division and square root
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed-FP32 instructions
so the kernels reproduce the reference arithmetic bit for bit where the tests
demand it.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsv.so")
ARCH = os.environ.get("SV_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["sv_project.hip", "sv_sort.hip", "sv_raster.hip", "sv_optim.hip", "sv_knn.hip", "sv_bin.hip", "sv_score.hip", "sv_noalloc.hip"]
HOST_SOURCES = ["sv_api.cpp", "sv_scene.cpp", "sv_ingest.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = [hipcc(), *HIP_FLAGS, *COMMON, "-c", path, "-o", out]
    else:
        cmd = [hipcc(), f"--offload-arch={ARCH}", *COMMON, "-pthread", "-c", path, "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"compile failed: {src}")
    return out


def build(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, HIP_SOURCES + HOST_SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    print(build(args.jobs))


if __name__ == "__main__":
    main()
