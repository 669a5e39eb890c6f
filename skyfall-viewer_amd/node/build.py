"""Build the Node N-API addon skyfall-viewer_amd/node/skyfall_sv.node over
libsv.so (host C++ only; the HIP code lives in libsv.so).

Usage: python skyfall-viewer_amd/node/build.py
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(HERE, "skyfall_sv.node")
NODE_INCLUDE = ("/usr/include/node", "/usr/include/nodejs/src")


def build() -> str | None:
    inc = next((d for d in NODE_INCLUDE if os.path.exists(os.path.join(d, "node_api.h"))), None)
    if inc is None:
        sys.stderr.write("node_api.h not found: the N-API addon is not built\n")
        return None
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", f"-I{inc}",
           os.path.join(HERE, "sv_napi.cpp"), "-o", OUT + ".tmp",
           f"-L{PKG}", "-lsv", "-Wl,-rpath,$ORIGIN/.."]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("N-API addon build failed")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build())
