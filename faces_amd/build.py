"""Builds the native parts of faces_amd in-tree.

  faces_amd/lib/libfaces_amd.so   HIP kernels + C ABI (include/faces_amd.h), gfx950

Run ``python faces_amd/build.py`` (or ``__graft_entry__.build()``).  The
command is one plain hipcc invocation; every HIP translation unit is built for
gfx950 only and without packed-FP32 instructions.
"""
from __future__ import annotations

import os
import shlex
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
INCLUDE = os.path.join(ROOT, "include")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["faces_plan.hip", "faces_flow.hip", "faces_flow_g0.hip", "faces_flow_g1.hip", "faces_flow_g2.hip",
               "faces_flow_g3.hip", "faces_batch.hip", "faces_serve.cpp", "faces_noalloc.hip"]
OBJDIR = os.path.join(PKG, "_build")

COMMON_FLAGS = [
    "-std=c++17",
    "-O3",
    "-fPIC",
    "-ffp-contract=off",
    "-Wall",
    "-Wno-unused-function",
    f"-I{INCLUDE}",
]
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    # Packed FP32 forms are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_library(verbose: bool = True) -> str:
    """Compiles every translation unit to an object (in parallel: the kernel
    instantiations of the plan and of the one-item kernel take minutes each),
    then links them into one shared library."""
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libfaces_amd.so")
    tmp = out + ".tmp"
    jobs = []
    for s in HIP_SOURCES:
        p = os.path.join(CSRC, s)
        if not os.path.exists(p):
            continue
        obj = os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o")
        src = ["-x", "hip", p] if s.endswith(".cpp") else [p]
        cmd = [HIPCC, *DEVICE_FLAGS, *COMMON_FLAGS, "-c", *src, "-o", obj]
        print("+", " ".join(shlex.quote(c) for c in cmd), flush=True)
        jobs.append((s, obj, subprocess.Popen(cmd)))
    failed = [s for s, _, proc in jobs if proc.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, f"hipcc -c {' '.join(failed)}")
    _run([HIPCC, *DEVICE_FLAGS, "-fPIC", "-shared", "-o", tmp, *[obj for _, obj, _ in jobs], "-lpthread"])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_library()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
