"""hrweno_hip -- MI355X (gfx950) HR-WENO hot path behind the reference's API.

See include/hrweno_capi.h for the C-ABI; this package is its Python host mirror.
"""
from ._lib import HrwError, LIB_PATH  # noqa: F401
from .api import (  # noqa: F401
    BC_INFLOW0_OUTFLOW, BC_PERIODIC, BC_ZERO_FLUX, BC_ZERO_GRADIENT, EXACT, FAST, FN_BURGERS,
    FN_LINEAR, GODUNOV, LAX_FRIEDRICHS, PBE2D, Adjoint, Burgers1D, BurgersBatch, Grid1, MsTvd, Problem, RkTvd, Weno,
    godunov, is_done, lax_friedrichs, tables, version,
)
