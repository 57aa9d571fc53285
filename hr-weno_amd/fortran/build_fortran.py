"""Compile the Fortran drop-in modules and the Fortran programs that use them.

    python hr-weno_amd/fortran/build_fortran.py

Outputs (in-tree, git-ignored): hr-weno_amd/fortran/_build/mod/*.mod, libhrweno_f.a, and the
executables in _build/bin (tests from tests/fortran/, examples from fortran/examples/), linked
against hr-weno_amd/lib/libhrweno_hip.so.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
REPO = os.path.dirname(PKG)
BUILD = os.path.join(HERE, "_build")
MOD = os.path.join(BUILD, "mod")
OBJ = os.path.join(BUILD, "obj")
BIN = os.path.join(BUILD, "bin")
LIBDIR = os.path.join(PKG, "lib")
FC = os.environ.get("FC_HRW", "/opt/rocm/bin/amdflang")
FFLAGS = ["-O2", "-ffp-contract=off", "-fPIC"]

# dependency order
MODULES = ["hrweno_kinds.F90", "hrweno_capi.f90", "hrweno_weno.f90", "hrweno_fluxes.f90",
           "hrweno_grids.f90", "hrweno_problems.f90", "hrweno_tvdode.f90"]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("Fortran compile failed")


def build() -> None:
    for d in (MOD, OBJ, BIN):
        os.makedirs(d, exist_ok=True)
    objs = []
    for m in MODULES:
        o = os.path.join(OBJ, os.path.splitext(m)[0] + ".o")
        run([FC, *FFLAGS, "-module-dir", MOD, "-c", os.path.join(HERE, m), "-o", o])
        objs.append(o)
    lib = os.path.join(BUILD, "libhrweno_f.a")
    if os.path.exists(lib):
        os.remove(lib)
    run(["ar", "rcs", lib, *objs])
    link = [lib, "-L" + LIBDIR, "-lhrweno_hip", "-Wl,-rpath," + LIBDIR,
            "-Wl,-rpath,$ORIGIN/../../../lib"]
    support = sorted(glob.glob(os.path.join(REPO, "tests", "fortran", "*_mod.f90")))
    sup_objs = []
    for s in support:
        o = os.path.join(OBJ, os.path.basename(s)[:-4] + ".o")
        run([FC, *FFLAGS, "-I", MOD, "-module-dir", MOD, "-c", s, "-o", o])
        sup_objs.append(o)
    progs = sorted(glob.glob(os.path.join(REPO, "tests", "fortran", "test_*.f90")))
    progs += sorted(glob.glob(os.path.join(HERE, "examples", "*.f90")))
    for p in progs:
        exe = os.path.join(BIN, os.path.basename(p)[:-4])
        run([FC, *FFLAGS, "-I", MOD, p, *sup_objs, "-o", exe, *link])


if __name__ == "__main__":
    build()
    print("built Fortran modules and programs in", BUILD)
