"""Builds libcig.so (HIP engine + host utilities) in-tree for gfx950.

Usage: python -m coincgraph_amd.build [--debug]
The library lands next to this file (coincgraph_amd/libcig.so) so that it
travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcig.so")
SAMPLES_BIN = os.path.join(HERE, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CIG_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-result",
]
SOURCES = ["engine.hip", "host_util.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_library(debug: bool = False) -> str:
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(build_dir, src + ".o")
        flags = list(HIP_FLAGS)
        if debug:
            flags = [f if f != "-O3" else "-O1" for f in flags] + ["-g"]
        if src.endswith(".cpp"):
            cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-Wall", "-c",
                   os.path.join(CSRC, src), "-o", obj, "-pthread"]
        else:
            cmd = [HIPCC] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-pthread"])
    os.replace(tmp, LIB)
    return LIB


def build_samples() -> list:
    """C++17 samples hybrid_{bfs,sssp,cc,pr}: one driver compiled per app."""
    os.makedirs(SAMPLES_BIN, exist_ok=True)
    src = os.path.join(HERE, "samples", "hybrid_main.cpp")
    if not os.path.exists(src):
        return []
    out = []
    for app in ("bfs", "sssp", "cc", "pr"):
        exe = os.path.join(SAMPLES_BIN, f"hybrid_{app}")
        _run(["g++", "-O2", "-std=c++17", "-Wall", f"-DCIG_SAMPLE_APP=\"{app}\"",
              "-I", os.path.join(ROOT, "include"), src, "-o", exe,
              "-L", HERE, "-lcig", f"-Wl,-rpath,$ORIGIN/..", "-pthread"])
        out.append(exe)
    return out


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise SystemExit("hipcc not found")
    build_library(debug="--debug" in argv)
    build_samples()


if __name__ == "__main__":
    main()
