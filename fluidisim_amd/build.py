"""Build recipe of libfsp.so (HIP kernels + C ABI for gfx950) in-tree.

hipcc compiles the device code for gfx950 with packed-FP32 instructions
disabled; the input generators are plain host C++ built with g++.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfsp.so")
ARCH = os.environ.get("FSP_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True, out: str = LIB, build_dir: str = BUILD, defines=()) -> str:
    """Compiles libfsp.so (to `out`, objects in `build_dir`; `defines` are
    extra -D macros for tuning builds)."""
    os.makedirs(build_dir, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    fsp_o = os.path.join(build_dir, "fsp.o")
    net_o = os.path.join(build_dir, "learned.o")
    gen_o = os.path.join(build_dir, "generate.o")
    dflags = [f"-D{d}" for d in defines]
    jobs = [
        [hipcc, *HIPCC_FLAGS, *dflags, "-c", os.path.join(CSRC, "fsp.hip"), "-o", fsp_o],
        [hipcc, *HIPCC_FLAGS, *dflags, "-c", os.path.join(CSRC, "learned.hip"), "-o", net_o],
        ["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-pthread", "-c",
         os.path.join(CSRC, "generate.cpp"), "-o", gen_o],
    ]
    procs = []
    for cmd in jobs:
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    failed = [cmd for cmd, p in zip(jobs, procs) if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    tmp = out + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-o", tmp, fsp_o, net_o, gen_o, "-pthread"])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
