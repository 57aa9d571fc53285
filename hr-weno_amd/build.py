"""Build the gfx950 shared library (and the Fortran drop-in modules) in-tree.

    python hr-weno_amd/build.py        # -> hr-weno_amd/lib/libhrweno_hip.so
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libhrweno_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: no implicit multiply-add fusion anywhere (the exact arithmetic form
# relies on it; the fast form spells out its fused multiply-adds).  Packed FP32 is disabled
# for all device code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def build_hip(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-o", tmp, os.path.join(CSRC, "capi.hip")]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l]
    if r.returncode != 0:
        sys.stderr.write("\n".join(errs) + "\n")
        raise RuntimeError("hipcc failed")
    if errs and verbose:
        sys.stderr.write("\n".join(errs) + "\n")
    os.replace(tmp, LIB)
    return LIB


def build(verbose: bool = False) -> None:
    build_hip(verbose)
    ftn = os.path.join(PKG, "fortran", "build_fortran.py")
    if os.path.exists(ftn):
        subprocess.run([sys.executable, ftn], check=True)


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
