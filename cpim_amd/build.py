"""Build the native libraries in-tree.

  libcpim_gpu.so       hipcc, gfx950: the propagation engine and its C-ABI
  libcpim_workload.so  g++: seeded synthetic instance generators

Host code is compiled for portable x86-64-v2 so that the libraries built in
one container run on any GPU host.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
GPU_LIB = os.path.join(HERE, "libcpim_gpu.so")
WORKLOAD_LIB = os.path.join(HERE, "libcpim_workload.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

GPU_SOURCES = ["engine.hip"]
GPU_DEPS = ["engine.hip", "batch.inc", "noalloc.inc", "kernels.h", "device.h"]

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay off in every HIP build of this project
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_gpu(verbose: bool = False) -> str:
    tmp = GPU_LIB + ".tmp"
    cmd = [HIPCC] + HIP_FLAGS + ["-I", os.path.join(REPO, "include"), "-o", tmp] + \
        [os.path.join(CSRC, s) for s in GPU_SOURCES]
    _run(cmd, verbose)
    os.replace(tmp, GPU_LIB)
    return GPU_LIB


def build_workload(verbose: bool = False) -> str:
    tmp = WORKLOAD_LIB + ".tmp"
    cmd = ["g++", "-O3", "-march=x86-64-v2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-o", tmp,
           os.path.join(CSRC, "workload.cpp"), "-lpthread"]
    _run(cmd, verbose)
    os.replace(tmp, WORKLOAD_LIB)
    return WORKLOAD_LIB


SOLVE_BIN = os.path.join(HERE, "cpim_solve")


def build_cli(verbose: bool = False) -> str:
    """cpim_solve: C++ driver over the facade, linked to both libraries."""
    tmp = SOLVE_BIN + ".tmp"
    cmd = ["g++", "-O2", "-march=x86-64-v2", "-std=c++17", "-Wall", "-o", tmp,
           os.path.join(CSRC, "cpim_solve.cpp"), "-I", os.path.join(REPO, "include"),
           "-L", HERE, "-lcpim_gpu", "-lcpim_workload", "-Wl,-rpath,$ORIGIN"]
    _run(cmd, verbose)
    os.replace(tmp, SOLVE_BIN)
    return SOLVE_BIN


def build_all(verbose: bool = False) -> None:
    build_workload(verbose)
    build_gpu(verbose)
    build_cli(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
