"""The Fortran drop-in modules (hr-weno_amd/fortran) run the reference's own test suites and the
example-1 composition.  The programs are built by __graft_entry__.build(); when they are
missing here (CPU container) they are compiled first."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "hr-weno_amd", "fortran", "_build", "bin")


def program(name):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        if not os.path.exists("/opt/rocm/bin/amdflang"):
            pytest.skip("Fortran programs are not built and amdflang is absent")
        subprocess.run(["python3", os.path.join(REPO, "hr-weno_amd", "fortran", "build_fortran.py")],
                       check=True)
    return exe


def run(name, timeout=120):
    r = subprocess.run([program(name)], capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-4000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failed" in r.stdout
    return r.stdout


def test_fortran_fluxes_and_grids_host():
    """test/test_fluxes.f90 and test/test_grid.f90 through the drop-in modules (host only)."""
    run("test_fluxes_grid")


@pytest.mark.gpu
def test_fortran_weno():
    """test/test_hrweno.f90 through the drop-in module, plus strided and device-array forms."""
    run("test_weno")


@pytest.mark.gpu
def test_fortran_tvdode():
    """test/test_tvdode.f90 through the drop-in module (device time loop, host callback)."""
    run("test_tvdode")


@pytest.mark.gpu
def test_fortran_dropin_example1():
    """example 1 as a user RHS on the drop-in API == the fused device problem, bit for bit."""
    run("test_dropin_burgers")


@pytest.mark.gpu
def test_fortran_adjoint():
    """rktvd%tape / %adjoint on a fused problem: FD agreement, conserved total, checkpoints."""
    run("test_adjoint")


@pytest.mark.gpu
def test_fortran_handles_device_integrand_state_lines():
    """Ownership of the drop-in integrators (1000 created and dropped with device memory flat,
    copies share, a temporary problem stays alive), rktvd_dev/mstvd_dev == fused, save/load
    resumes bit for bit, reconstruct_lines == reconstruct line by line (example 2's loops)."""
    run("test_handles", timeout=300)
