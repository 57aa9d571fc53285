"""ctypes access to the CPU checker (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module; the shipped library never does.

* ``Oracle``    -- the C restatement of the reference arithmetic (``hrweno_oracle.c``).
* ``Reference`` -- the reference's own ``hrweno_weno``/``hrweno_fluxes`` modules compiled
  from its sources into ``oracle/_ref`` (absent when the reference checkout was not there
  when the tree was built).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "libhrweno_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libhrweno_ref.so")

_P = C.POINTER(C.c_double)
_I = C.c_int
_L = C.c_int64
_D = C.c_double

FLUX_LF, FLUX_GODUNOV = 0, 1
BC_PERIODIC, BC_ZERO_GRADIENT = 0, 1


def build(ref: bool = True) -> None:
    """Compile the checker (and the reference build when its sources are present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """The C restatement.  Every method returns fresh numpy arrays."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build(ref=False)
        lib = C.CDLL(path)
        lib.hro_weno_reconstruct.argtypes = [_I, _I, _D, _P, _P, _L, _P, _L, _P, _L]
        lib.hro_weno_calc_cnu.argtypes = [_I, _I, _P, _P]
        lib.hro_weno_tables.argtypes = [_I, _P, _P]
        lib.hro_lax_friedrichs.argtypes = [_D, _D, _D, _D, _D]
        lib.hro_lax_friedrichs.restype = _D
        lib.hro_godunov.argtypes = [_D, _D, _D, _D]
        lib.hro_godunov.restype = _D
        lib.hro_grid_linear.argtypes = [_D, _D, _I, _P]
        lib.hro_grid_bilinear.argtypes = [_D, _D, _D, _I, _I, _P]
        lib.hro_grid_log.argtypes = [_D, _D, _I, _P]
        lib.hro_grid_geometric.argtypes = [_D, _D, _D, _I, _P]
        lib.hro_grid_compute.argtypes = [_I, _P, _P, _P]
        lib.hro_burgers_rk.argtypes = [_I, _I, _D, _I, _I, _D, _P, _P, _I, _P, _P, _D, _D,
                                       _I, C.POINTER(_I)]
        lib.hro_burgers_rk.restype = _L
        lib.hro_pbe_ms.argtypes = [_I, _I, _I, _D, _I, _D, _D, _D, _D, _P, _P, _I, _P, _P, _D,
                                   _D]
        lib.hro_pbe_rhs.argtypes = [_I, _I, _I, _D, _I, _D, _D, _D, _D, _P, _P, _P, _P]
        lib.hro_pbe_plant.argtypes = [_I]
        lib.hro_burgers_ms.argtypes = [_I, _I, _D, _I, _I, _D, _P, _P, _P, _P, _D, _D]
        lib.hro_linear_ode.argtypes = [_I, _P, _I, _P, _P, _D, _D]
        lib.hro_burgers_rhs.argtypes = [_I, _I, _D, _I, _I, _D, _P, _P, _P]
        lib.hro_burgers_rhs_vjp.argtypes = [_I, _I, _D, _D, _P, _P, _P, _P]
        lib.hro_burgers_rk_adjoint.argtypes = [_I, _I, _D, _D, _P, _I, _L, _D, _P, _P]
        lib.hro_burgers_rhs_vjp_ex.argtypes = [_I, _I, _D, _I, _I, _D, _P, _P, _P, _P]
        lib.hro_burgers_rk_adjoint_ex.argtypes = [_I, _I, _D, _I, _I, _D, _P, _I, _L, _D, _P, _P]
        lib.hro_burgers_ms_adjoint.argtypes = [_I, _I, _D, _I, _I, _D, _P, _P, _D, _D, _P, _P,
                                               C.POINTER(C.c_int64)]
        lib.hro_adjoint_plant.argtypes = [_I]
        lib.hro_set_threads.argtypes = [_I]
        lib.hro_get_threads.restype = _I
        self.lib = lib

    # -- threads ---------------------------------------------------------------------
    def set_threads(self, n: int) -> None:
        self.lib.hro_set_threads(int(n))

    # -- weno ------------------------------------------------------------------------
    def tables(self, k: int):
        d = np.zeros(3)
        c = np.zeros(12)
        self.lib.hro_weno_tables(k, _ptr(d), _ptr(c))
        return d[:k].copy(), c[: k * (k + 1)].reshape(k + 1, k).copy()  # c[r+1, j]

    def calc_cnu(self, k: int, xedges) -> np.ndarray:
        xe = _f64(xedges)
        n = xe.size - 1
        cnu = np.zeros(n * k * (k + 1))
        rc = self.lib.hro_weno_calc_cnu(n, k, _ptr(xe), _ptr(cnu))
        if rc:
            raise ValueError(f"hro_weno_calc_cnu failed ({rc})")
        return cnu.reshape(n, k + 1, k)  # [i-1, r+1, j]

    def reconstruct(self, v, k: int = 3, eps: float = 1e-6, cnu=None):
        v = _f64(v)
        n = v.size
        vl = np.empty(n)
        vr = np.empty(n)
        cp = _ptr(_f64(cnu)) if cnu is not None else None
        keep = _f64(cnu) if cnu is not None else None
        if keep is not None:
            cp = _ptr(keep)
        rc = self.lib.hro_weno_reconstruct(n, k, eps, cp, _ptr(v), 1, _ptr(vl), 1, _ptr(vr), 1)
        if rc:
            raise ValueError(f"hro_weno_reconstruct failed ({rc})")
        return vl, vr

    # -- fluxes ----------------------------------------------------------------------
    def lax_friedrichs(self, fm, fp, vm, vp, alpha) -> float:
        return self.lib.hro_lax_friedrichs(fm, fp, vm, vp, alpha)

    def godunov(self, fm, fp, vm, vp) -> float:
        return self.lib.hro_godunov(fm, fp, vm, vp)

    # -- grids -----------------------------------------------------------------------
    def grid_linear(self, xmin, xmax, n):
        e = np.zeros(n + 1)
        if self.lib.hro_grid_linear(xmin, xmax, n, _ptr(e)):
            raise ValueError("invalid grid")
        return e

    def grid_bilinear(self, xmin, xcross, xmax, n1, n2):
        e = np.zeros(n1 + n2 + 1)
        if self.lib.hro_grid_bilinear(xmin, xcross, xmax, n1, n2, _ptr(e)):
            raise ValueError("invalid grid")
        return e

    def grid_log(self, xmin, xmax, n):
        e = np.zeros(n + 1)
        if self.lib.hro_grid_log(xmin, xmax, n, _ptr(e)):
            raise ValueError("invalid grid")
        return e

    def grid_geometric(self, xmin, xmax, ratio, n):
        e = np.zeros(n + 1)
        if self.lib.hro_grid_geometric(xmin, xmax, ratio, n, _ptr(e)):
            raise ValueError("invalid grid")
        return e

    def grid_compute(self, edges):
        e = _f64(edges)
        n = e.size - 1
        c = np.zeros(n)
        w = np.zeros(n)
        self.lib.hro_grid_compute(n, _ptr(e), _ptr(c), _ptr(w))
        return c, w

    # -- integrators -----------------------------------------------------------------
    def burgers_rk(self, u0, width, *, k=3, eps=1e-6, flux=FLUX_LF, bc=BC_PERIODIC,
                   alpha=-1.0, order=3, t0=0.0, tout=1.0, dt=1e-3, itask=1, cnu=None):
        """Advance 1D Burgers with rktvd; returns (u, t, steps, fevals)."""
        u = _f64(u0).copy()
        w = _f64(width)
        t = C.c_double(t0)
        fe = C.c_int(0)
        cn = _f64(cnu) if cnu is not None else None
        steps = self.lib.hro_burgers_rk(u.size, k, eps, flux, bc, alpha, _ptr(w),
                                        _ptr(cn) if cn is not None else None, order,
                                        _ptr(u), C.byref(t), tout, dt, itask, C.byref(fe))
        if steps < 0:
            raise ValueError("hro_burgers_rk failed")
        return u, t.value, int(steps), fe.value

    def burgers_rhs(self, v, width, *, k=3, eps=1e-6, flux=FLUX_LF, bc=BC_PERIODIC, alpha=-1.0):
        """One evaluation of the Burgers RHS L(v)."""
        v, w = _f64(v), _f64(width)
        out = np.empty(v.size)
        if self.lib.hro_burgers_rhs(v.size, k, eps, flux, bc, alpha, _ptr(w), _ptr(v), _ptr(out)):
            raise ValueError("hro_burgers_rhs failed")
        return out

    # -- discrete adjoint (Burgers: Lax-Friedrichs or Godunov, periodic or zero-gradient) --
    def burgers_rhs_vjp(self, v, mu, width, *, k=3, eps=1e-6, alpha=-1.0, flux=FLUX_LF,
                        bc=BC_PERIODIC):
        """(dL/dv)^T mu for the Burgers RHS."""
        v, mu, w = _f64(v), _f64(mu), _f64(width)
        out = np.empty(v.size)
        if self.lib.hro_burgers_rhs_vjp_ex(v.size, k, eps, flux, bc, alpha, _ptr(w), _ptr(v),
                                           _ptr(mu), _ptr(out)):
            raise ValueError("hro_burgers_rhs_vjp_ex failed")
        return out

    def burgers_rk_adjoint(self, u0, gN, width, *, nsteps, dt, k=3, eps=1e-6, alpha=-1.0,
                           order=3, flux=FLUX_LF, bc=BC_PERIODIC):
        """nsteps rktvd steps from u0 and the reverse sweep of gN; returns (u_N, dJ/du0)."""
        u = _f64(u0).copy()
        g = _f64(gN).copy()
        w = _f64(width)
        if self.lib.hro_burgers_rk_adjoint_ex(u.size, k, eps, flux, bc, alpha, _ptr(w), order,
                                              int(nsteps), dt, _ptr(u), _ptr(g)):
            raise ValueError("hro_burgers_rk_adjoint_ex failed")
        return u, g

    def burgers_ms_adjoint(self, u0, gN, width, *, t0=0.0, tout, dt, k=3, eps=1e-6, alpha=-1.0,
                           flux=FLUX_LF, bc=BC_PERIODIC):
        """mstvd from a fresh object (t0 -> tout) and the reverse sweep of gN; returns
        (u_N, dJ/du0, t, N) with N the steps taken, the four starting steps included."""
        u = _f64(u0).copy()
        g = _f64(gN).copy()
        w = _f64(width)
        t = C.c_double(t0)
        ns = C.c_int64(0)
        if self.lib.hro_burgers_ms_adjoint(u.size, k, eps, flux, bc, alpha, _ptr(w), C.byref(t),
                                           tout, dt, _ptr(u), _ptr(g), C.byref(ns)):
            raise ValueError("hro_burgers_ms_adjoint failed")
        return u, g, t.value, int(ns.value)

    def adjoint_plant(self, which: int) -> None:
        """Test hook: plant mistake `which` (1..8) in the adjoint, 0 restores it."""
        self.lib.hro_adjoint_plant(int(which))

    def pbe_ms(self, u0, edges1, edges2, *, k=3, eps=1e-6, bc=1, g1=(1.0, 0.5), g2=(0.5, 0.0),
               t0=0.0, tout=1.0, dt=1e-3, order=0):
        """Advance the 2D PBE from a fresh object with mstvd (order=0) or rktvd(order);
        returns (u, t, fevals)."""
        u = _f64(u0).copy()
        e1, e2 = _f64(edges1), _f64(edges2)
        t = C.c_double(t0)
        fe = self.lib.hro_pbe_ms(e1.size - 1, e2.size - 1, k, eps, bc, g1[0], g1[1], g2[0],
                                 g2[1], _ptr(e1), _ptr(e2), order, _ptr(u), C.byref(t), tout,
                                 dt)
        return u, t.value, fe

    def pbe_rhs(self, v, edges1, edges2, *, k=3, eps=1e-6, bc=1, g1=(1.0, 0.5), g2=(0.5, 0.0)):
        """One evaluation of the 2D PBE right-hand side (example 2's composition)."""
        v = _f64(v)
        e1, e2 = _f64(edges1), _f64(edges2)
        out = np.empty_like(v)
        if self.lib.hro_pbe_rhs(e1.size - 1, e2.size - 1, k, eps, bc, g1[0], g1[1], g2[0],
                                g2[1], _ptr(e1), _ptr(e2), _ptr(v), _ptr(out)) != 0:
            raise MemoryError("hro_pbe_rhs")
        return out

    def pbe_plant(self, which: int) -> None:
        """Test hook: plant mistake `which` (1..3) in the 2D composition, 0 restores it."""
        self.lib.hro_pbe_plant(int(which))

    def burgers_ms(self, u0, width, *, k=3, eps=1e-6, flux=FLUX_LF, bc=BC_PERIODIC, alpha=-1.0,
                   t0=0.0, tout=1.0, dt=1e-3, cnu=None):
        """Advance 1D Burgers with mstvd from a fresh object; returns (u, t, fevals)."""
        u = _f64(u0).copy()
        w = _f64(width)
        cn = _f64(cnu) if cnu is not None else None
        t = C.c_double(t0)
        fe = self.lib.hro_burgers_ms(u.size, k, eps, flux, bc, alpha, _ptr(w),
                                     _ptr(cn) if cn is not None else None, _ptr(u),
                                     C.byref(t), tout, dt)
        return u, t.value, fe

    def linear_ode(self, a, u0, order, t0, tout, dt):
        """u' = a*u with rktvd(order) or mstvd (order=0); returns (u, t, fevals)."""
        a = _f64(a)
        u = _f64(u0).copy()
        t = C.c_double(t0)
        fe = self.lib.hro_linear_ode(a.size, _ptr(a), order, _ptr(u), C.byref(t), tout, dt)
        return u, t.value, fe


class Reference:
    """The reference's own hrweno_weno / hrweno_fluxes, compiled from its sources."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        lib = C.CDLL(path)
        lib.ref_weno_reconstruct.argtypes = [_I, _I, _D, _I, _P, _P, _P, _P]
        lib.ref_weno_cnu.argtypes = [_I, _I, _P, _P]
        lib.ref_weno_reconstruct_repeat.argtypes = [_I, _I, _D, _I, _P, _P, _P]
        lib.ref_fluxes.argtypes = [_I, _I, _I, _P, _P, _P, _D, _D, _P]
        lib.ref_burgers_rk3.argtypes = [_I, _I, _D, _I, _D, _P, _P]
        lib.ref_burgers_ex.argtypes = [_I, _I, _D, _I, _I, _I, _I, _D, _P, _P]
        self.lib = lib

    def burgers_rk3(self, u0, width, nsteps, dt, k=3, eps=1e-6):
        """nsteps periodic WENO-LF-RK3 steps through the reference's own modules."""
        u = _f64(u0).copy()
        w = _f64(width)
        self.lib.ref_burgers_rk3(u.size, k, eps, nsteps, dt, _ptr(w), _ptr(u))
        return u

    def burgers_ex(self, u0, width, nsteps, dt, *, k=3, eps=1e-6, flux=FLUX_LF,
                   bc=BC_PERIODIC, method=3):
        """Example 1's composition on the reference's own modules: nsteps RK3 steps
        (method 3) or the multistep method with its four RK3 starting steps (method 0)."""
        u = _f64(u0).copy()
        w = _f64(width)
        self.lib.ref_burgers_ex(u.size, k, eps, flux, bc, method, int(nsteps), dt, _ptr(w),
                                _ptr(u))
        return u

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_SO)

    def reconstruct(self, v, k=3, eps=1e-6, xedges=None):
        v = _f64(v)
        n = v.size
        vl = np.empty(n)
        vr = np.empty(n)
        xe = _f64(xedges) if xedges is not None else np.zeros(n + 1)
        self.lib.ref_weno_reconstruct(n, k, eps, int(xedges is not None), _ptr(xe), _ptr(v),
                                      _ptr(vl), _ptr(vr))
        return vl, vr

    def cnu(self, k, xedges):
        xe = _f64(xedges)
        n = xe.size - 1
        out = np.zeros(n * k * (k + 1))
        self.lib.ref_weno_cnu(n, k, _ptr(xe), _ptr(out))
        return out.reshape(n, k + 1, k)

    def reconstruct_repeat(self, v, k, eps, nrep):
        v = _f64(v)
        vl = np.empty(v.size)
        vr = np.empty(v.size)
        self.lib.ref_weno_reconstruct_repeat(v.size, k, eps, nrep, _ptr(v), _ptr(vl), _ptr(vr))
        return vl, vr

    def fluxes(self, kind, which, vm, vp, x, t, alpha=0.0):
        vm, vp, x = _f64(vm), _f64(vp), _f64(x)
        h = np.empty(vm.size)
        self.lib.ref_fluxes(kind, which, vm.size, _ptr(vm), _ptr(vp), _ptr(x), t, alpha, _ptr(h))
        return h
