"""ctypes binding of libhrweno_hip.so (the C-ABI in include/hrweno_capi.h).

The library must have been built in-tree (``python hr-weno_amd/build.py``).  There is no
fallback: if the shared library or a GPU is missing, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libhrweno_hip.so")

P = C.c_void_p      # device pointers and handles
D = C.c_double
I = C.c_int
L = C.c_int64
PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)
PL = C.POINTER(C.c_int64)

INTEGRAND = C.CFUNCTYPE(C.c_int, P, D, P, P, P)

# name -> argtypes (every symbol declared in include/hrweno_capi.h)
SIGNATURES = {
    "hrw_version": [],
    "hrw_last_error": [],
    "hrw_last_error_copy": [C.c_char_p, L],
    "hrw_device_count": [PI],
    "hrw_device_alloc": [C.c_size_t, C.POINTER(P)],
    "hrw_device_free": [P],
    "hrw_device_mem_info": [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "hrw_memcpy_htod": [P, P, C.c_size_t, P],
    "hrw_memcpy_dtoh": [P, P, C.c_size_t, P],
    "hrw_memcpy_dtod": [P, P, C.c_size_t, P],
    "hrw_stream_create": [C.POINTER(P)],
    "hrw_stream_destroy": [P],
    "hrw_stream_synchronize": [P],
    "hrw_timer_start": [P],
    "hrw_timer_stop": [P, PD],
    "hrw_weno_tables": [I, PD, PD],
    "hrw_weno_create": [I, I, D, PD, I, C.POINTER(P)],
    "hrw_weno_destroy": [P],
    "hrw_weno_info": [P, PI, PI, PD, PI],
    "hrw_weno_get_cnu": [P, PD],
    "hrw_weno_reconstruct": [P, P, L, P, L, P, L, P],
    "hrw_weno_reconstruct_lines": [P, L, P, L, L, P, P, L, L, P],
    "hrw_weno_reconstruct_host": [P, PD, L, PD, PD],
    "hrw_weno_calc_cnu": [I, I, PD, PD],
    "hrw_reconstruct_host": [I, I, D, I, PD, PD, L, PD, L, PD, L],
    "hrw_reconstruct_dev": [I, I, D, I, P, L, P, L, P, L, P],
    "hrw_reconstruct_lines_host": [I, I, D, I, PD, L, PD, L, L, PD, PD, L, L],
    "hrw_reconstruct_lines_dev": [I, I, D, I, L, P, L, L, P, P, L, L, P],
    "hrw_flux_lax_friedrichs": [I, D, D, L, P, P, P, D, D, P, P],
    "hrw_flux_godunov": [I, D, D, L, P, P, P, D, P, P],
    "hrw_grid_linear": [D, D, I, PD],
    "hrw_grid_bilinear": [D, D, D, I, I, PD],
    "hrw_grid_log": [D, D, I, PD],
    "hrw_grid_geometric": [D, D, D, I, PD],
    "hrw_grid_compute": [I, PD, PD, PD],
    "hrw_problem_burgers1d": [I, I, D, I, I, D, D, D, PD, I, C.POINTER(P)],
    "hrw_problem_burgers1d_edges": [I, I, D, I, I, D, PD, I, C.POINTER(P)],
    "hrw_problem_pbe2d": [I, I, I, D, I, D, D, D, D, D, D, D, D, I, C.POINTER(P)],
    "hrw_problem_tune": [P, I, I],
    "hrw_problem_destroy": [P],
    "hrw_problem_neq": [P, PL],
    "hrw_problem_rhs": [P, D, P, P, P],
    "hrw_rktvd_create": [P, I, C.POINTER(P)],
    "hrw_rktvd_create_cb": [INTEGRAND, P, L, I, I, C.POINTER(P)],
    "hrw_mstvd_create": [P, C.POINTER(P)],
    "hrw_mstvd_create_cb": [INTEGRAND, P, L, I, C.POINTER(P)],
    "hrw_ode_destroy": [P],
    "hrw_ode_set_stream": [P, P],
    "hrw_ode_integrate": [P, P, PD, D, D, I],
    "hrw_ode_info": [P, PL, PI, PI, PI],
    "hrw_ode_set_check": [P, I],
    "hrw_ode_state_size": [P, PL],
    "hrw_ode_save": [P, P, D, P, L],
    "hrw_ode_load": [P, P, PD, P, L],
    "hrw_rktvd_steps_noalloc": [P, I, P, L, D, P],
    "hrw_batch_burgers_create": [L, PI, PI, PD, PD, PD, D, I, D, I, C.POINTER(P)],
    "hrw_batch_destroy": [P],
    "hrw_batch_size": [P, PL, PL],
    "hrw_batch_set_stream": [P, P],
    "hrw_batch_synchronize": [P],
    "hrw_batch_step": [P, P, L, PD, I],
    "hrw_batch_step_per_line": [P, P, L, PD],
    "hrw_batch_tune": [P, P, PI],
    "hrw_batch_set_cells_per_lane": [P, I],
    "hrw_batch_set_form": [P, I],
    "hrw_batch_config": [P, PI, PI],
    "hrw_adjoint_create": [P, I, D, L, I, C.POINTER(P)],
    "hrw_batch_adjoint_create": [P, L, I, C.POINTER(P)],
    "hrw_adjoint_destroy": [P],
    "hrw_adjoint_set_stream": [P, P],
    "hrw_adjoint_set_cells_per_lane": [P, I],
    "hrw_adjoint_forward": [P, P, P],
    "hrw_adjoint_backward": [P, P, P],
    "hrw_adjoint_info": [P, PL, PL, PI, PL],
    "hrw_ode_tape": [P, I],
    "hrw_ode_adjoint": [P, P],
}

_lib = None


class HrwError(RuntimeError):
    """A non-zero hrw_status; the message is the library's (the reference's error texts)."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python hr-weno_amd/build.py` "
                "(there is no CPU fallback)")
        h = C.CDLL(LIB_PATH)
        for name, argtypes in SIGNATURES.items():
            f = getattr(h, name)
            f.argtypes = argtypes
            f.restype = C.c_int
        h.hrw_version.restype = C.c_char_p
        h.hrw_last_error.restype = C.c_char_p
        _lib = h
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().hrw_last_error().decode()
        raise HrwError(rc, msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
