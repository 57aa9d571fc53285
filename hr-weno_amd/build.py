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
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
OBJDIR = os.path.join(PKG, "_build", "obj")
# translation units (source, extra flags, object): the launchers instantiate most of the device
# code, so they are compiled apart and in parallel
UNITS = [
    ("capi.hip", [], "capi.o"),
    ("launch_fast.hip", ["-DHRW_K=1"], "launch_fast_k1.o"),
    ("launch_fast.hip", ["-DHRW_K=2"], "launch_fast_k2.o"),
    ("launch_fast.hip", ["-DHRW_K=3"], "launch_fast_k3.o"),
    ("launch_kernels.hip", [], "launch_kernels.o"),
    ("launch_adjoint.hip", [], "launch_adjoint.o"),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l]
    if r.returncode != 0:
        sys.stderr.write("\n".join(errs) + "\n")
        raise RuntimeError("hipcc failed: " + " ".join(cmd))
    if errs and verbose:
        sys.stderr.write("\n".join(errs) + "\n")


def build_hip(verbose: bool = False) -> str:
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = [[HIPCC, *HIP_FLAGS, *extra, "-c", "-o", os.path.join(OBJDIR, obj), os.path.join(CSRC, src)]
            for src, extra, obj in UNITS]
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(workers) as ex:
        for f in [ex.submit(_run, j, verbose) for j in jobs]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp,
          *[os.path.join(OBJDIR, obj) for _, _, obj in UNITS]], verbose)
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
