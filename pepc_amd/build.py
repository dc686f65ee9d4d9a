"""Builds the native parts in-tree.

* ``pepc_amd/libpepc_gpu.so``: the HIP kernels and the C-ABI of
  ``include/pepc_gpu.h`` (hipcc, gfx950 only).
* ``pepc_amd/fortran/*``: the iso_c_binding shim and the pepc-mini style
  driver (amdflang), when amdflang is present.

Usage: ``python -m pepc_amd.build [--no-fortran]``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpepc_gpu.so")
ARCH = os.environ.get("PEPC_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["keys.hip", "sort.hip", "tree.hip", "walk.hip", "step.hip", "adjoint.hip",
               "lattice.hip", "runtime.hip"]

# -ffp-contract=off: the key, box, moment and acceptance arithmetic must run in
# the written operation order (the FP32 kernels request fused multiply-adds
# explicitly).  Packed FP32 instructions are disabled on this target.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_hip(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = LIB + ".tmp"
    _run([hipcc(), *HIP_FLAGS, "-o", tmp, *srcs])
    os.replace(tmp, LIB)
    return LIB


def build_fortran():
    fdir = os.path.join(HERE, "fortran")
    mk = os.path.join(fdir, "Makefile")
    if not os.path.exists(mk):
        return None
    flang = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(flang):
        print("amdflang not found; skipping the Fortran shim", flush=True)
        return None
    _run(["make", "-C", fdir, f"FC={flang}"])
    return fdir


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    build_hip()
    if "--no-fortran" not in argv:
        build_fortran()


if __name__ == "__main__":
    main()
