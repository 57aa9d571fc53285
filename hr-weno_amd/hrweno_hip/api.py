"""Python mirror of the reference's Fortran API over the gfx950 C-ABI.

Names, argument meanings and error texts follow the reference modules
(src/hrweno_weno.f90, src/hrweno_fluxes.f90, src/hrweno_grids.f90, src/hrweno_tvdode.f90).
Device arrays are torch float64 CUDA tensors (any 1D stride for ``reconstruct``); work is
enqueued on torch's current stream, so torch events time it.
"""
from __future__ import annotations

import ctypes as C
from typing import Callable, Optional

import numpy as np

from . import _lib
from ._lib import HrwError, call

EXACT, FAST = 0, 1
LAX_FRIEDRICHS, GODUNOV = 0, 1
FN_BURGERS, FN_LINEAR = 0, 1
BC_PERIODIC, BC_ZERO_GRADIENT, BC_ZERO_FLUX, BC_INFLOW0_OUTFLOW = 0, 1, 2, 3

_ARITH = {"exact": EXACT, "fast": FAST, EXACT: EXACT, FAST: FAST}


def _torch():
    import torch
    return torch


def _stream():
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dptr(t) -> C.c_void_p:
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64:
        raise TypeError("expected a float64 CUDA tensor")
    return C.c_void_p(t.data_ptr())


def _dvec(t, n: int, name: str, device=None, strided: bool = False) -> C.c_void_p:
    """Device pointer of a float64 CUDA tensor of exactly n elements, checked before any
    launch: contiguous (or, with strided=True, a 1D view of any positive stride), on `device`
    when given.  A short or strided array would otherwise become out-of-bounds or wrong device
    accesses."""
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64:
        raise TypeError(f"{name}: expected a float64 CUDA tensor")
    if t.numel() != n:
        raise HrwError(1, f"{name}: expected {n} values, got {t.numel()}")
    if strided:
        if t.dim() != 1 or (n > 1 and t.stride(0) < 1):
            raise HrwError(1, f"{name}: expected a 1D view with a positive stride")
    elif not t.is_contiguous():
        raise HrwError(1, f"{name}: expected a contiguous tensor")
    if device is not None and t.device != device:
        raise HrwError(1, f"{name}: on {t.device}, expected {device}")
    return C.c_void_p(t.data_ptr())


def _np(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _pd(a: np.ndarray):
    return a.ctypes.data_as(_lib.PD)


def _release(obj, destroy: str) -> None:
    """Destroy a library object once; silent while the interpreter shuts down."""
    h = getattr(obj, "_h", None)
    if not h:
        return
    obj._h = None
    try:
        getattr(_lib.lib(), destroy)(h)
    except Exception:
        pass


def version() -> str:
    return _lib.lib().hrw_version().decode()


# ---------------------------------------------------------------------------------- weno
def tables(k: int):
    """d(0:k-1) and c(0:k-1,-1:k-1) of src/hrweno_weno.f90:12-21; c is returned as c[j, r+1]."""
    d = np.zeros(3)
    c = np.zeros(12)
    call("hrw_weno_tables", k, _pd(d), _pd(c))
    return d[:k].copy(), c[: k * (k + 1)].reshape(k + 1, k).T.copy()


class Weno:
    """``weno(ncells, k, eps, xedges)`` (src/hrweno_weno.f90:23-127)."""

    def __init__(self, ncells: int, k: int = 3, eps: float = 1e-6, xedges=None,
                 arith="fast"):
        h = C.c_void_p()
        xe = None
        if xedges is not None:
            xe = _np(xedges)
            if xe.size != ncells + 1:
                raise HrwError(1, "Invalid input 'xedges': size(xedges) /= ncells + 1.")
        call("hrw_weno_create", int(ncells), int(k), float(eps),
             _pd(xe) if xe is not None else None, _ARITH[arith], C.byref(h))
        self._h = h
        self.ncells, self.k, self.eps = int(ncells), int(k), float(eps)
        self.uniform_grid = xe is None

    def __del__(self):
        _release(self, "hrw_weno_destroy")

    @property
    def cnu(self) -> np.ndarray:
        """cnu(0:k-1,-1:k-1,1:ncells) as an array indexed [i-1, r+1, j]."""
        k, n = self.k, self.ncells
        out = np.zeros(n * k * (k + 1))
        call("hrw_weno_get_cnu", self._h, _pd(out))
        return out.reshape(n, k + 1, k)

    def reconstruct(self, v, vl=None, vr=None):
        """``call w%reconstruct(v, vl, vr)`` (src/hrweno_weno.f90:129-219) on device tensors."""
        torch = _torch()
        if v.dim() != 1 or v.numel() != self.ncells:
            raise HrwError(1, "reconstruct: v must be a vector of ncells values")
        if vl is None:
            vl = torch.empty(self.ncells, dtype=torch.float64, device=v.device)
        if vr is None:
            vr = torch.empty(self.ncells, dtype=torch.float64, device=v.device)
        pv = _dvec(v, self.ncells, "v", strided=True)
        pl = _dvec(vl, self.ncells, "vl", v.device, strided=True)
        pr = _dvec(vr, self.ncells, "vr", v.device, strided=True)
        call("hrw_weno_reconstruct", self._h, pv, v.stride(0), pl, vl.stride(0), pr,
             vr.stride(0), _stream())
        return vl, vr

    def reconstruct_lines(self, v, nlines: int, incv: int, ldv: int, vl, vr, inco: int, ldo: int):
        """All nlines lines of a 2D array in one launch (element i of line l at l*ld + i*inc).
        The arrays are contiguous buffers that must hold the last element addressed."""
        n, L = self.ncells, int(nlines)
        if min(int(incv), int(inco)) < 1 or min(int(ldv), int(ldo)) < 0 or L < 1:
            raise HrwError(1, "reconstruct_lines: strides must be positive")
        need_v = (L - 1) * int(ldv) + (n - 1) * int(incv) + 1
        need_o = (L - 1) * int(ldo) + (n - 1) * int(inco) + 1
        for name, t, need in (("v", v, need_v), ("vl", vl, need_o), ("vr", vr, need_o)):
            _dptr(t)
            if not t.is_contiguous() or t.numel() < need:
                raise HrwError(1, f"reconstruct_lines: {name} must be a contiguous buffer of at "
                                  f"least {need} values")
            if t.device != v.device:
                raise HrwError(1, f"reconstruct_lines: {name} on {t.device}, v on {v.device}")
        call("hrw_weno_reconstruct_lines", self._h, L, _dptr(v), int(incv), int(ldv),
             _dptr(vl), _dptr(vr), int(inco), int(ldo), _stream())
        return vl, vr

    def reconstruct_host(self, v):
        """Host numpy arrays in and out (synchronous, through device staging buffers)."""
        v = np.asarray(v, dtype=np.float64)
        if v.ndim != 1 or v.size != self.ncells:
            raise HrwError(1, "reconstruct: v must be a vector of ncells values")
        step = v.strides[0] // 8
        base = v if step > 0 else np.ascontiguousarray(v)
        if step <= 0:
            step = 1
        vl = np.empty(self.ncells)
        vr = np.empty(self.ncells)
        call("hrw_weno_reconstruct_host", self._h, base.ctypes.data_as(_lib.PD), step, _pd(vl),
             _pd(vr))
        return vl, vr


# ---------------------------------------------------------------------------------- fluxes
def lax_friedrichs(vm, vp, x=None, t=0.0, alpha=1.0, fn=FN_BURGERS, coef=(0.0, 1.0), out=None):
    """Elementwise src/hrweno_fluxes.f90:22-45 for a built-in flux function."""
    torch = _torch()
    if out is None:
        out = torch.empty_like(vm)
    n = vm.numel()
    call("hrw_flux_lax_friedrichs", fn, float(coef[0]), float(coef[1]), n, _dvec(vm, n, "vm"),
         _dvec(vp, n, "vp", vm.device), _dvec(x, n, "x", vm.device) if x is not None else None,
         float(t), float(alpha), _dvec(out, n, "out", vm.device), _stream())
    return out


def godunov(vm, vp, x=None, t=0.0, fn=FN_BURGERS, coef=(0.0, 1.0), out=None):
    """Elementwise src/hrweno_fluxes.f90:47-76 for a built-in flux function."""
    torch = _torch()
    if out is None:
        out = torch.empty_like(vm)
    n = vm.numel()
    call("hrw_flux_godunov", fn, float(coef[0]), float(coef[1]), n, _dvec(vm, n, "vm"),
         _dvec(vp, n, "vp", vm.device), _dvec(x, n, "x", vm.device) if x is not None else None,
         float(t), _dvec(out, n, "out", vm.device), _stream())
    return out


# ---------------------------------------------------------------------------------- grids
class Grid1:
    """``type(grid1)`` (src/hrweno_grids.f90:12-37); arrays are host numpy."""

    def __init__(self):
        self.name = ""
        self.scale = ""
        self.ncells = 0
        self.edges = np.zeros(0)
        self.center = np.zeros(0)
        self.width = np.zeros(0)

    @property
    def left(self):
        return self.edges[:-1]

    @property
    def right(self):
        return self.edges[1:]

    def _compute(self, edges, name):
        n = edges.size - 1
        self.ncells = n
        self.edges = edges
        self.center = np.zeros(n)
        self.width = np.zeros(n)
        call("hrw_grid_compute", n, _pd(edges), _pd(self.center), _pd(self.width))
        self.name = name

    def linear(self, xmin, xmax, ncells, name=""):
        e = np.zeros(int(ncells) + 1) if ncells >= 1 else np.zeros(1)
        call("hrw_grid_linear", float(xmin), float(xmax), int(ncells), _pd(e))
        self.scale = "linear"
        self._compute(e, name)
        return self

    def bilinear(self, xmin, xcross, xmax, ncells, name=""):
        n1, n2 = int(ncells[0]), int(ncells[1])
        e = np.zeros(max(n1 + n2, 0) + 1)
        call("hrw_grid_bilinear", float(xmin), float(xcross), float(xmax), n1, n2, _pd(e))
        self.scale = "bilinear"
        self._compute(e, name)
        return self

    def log(self, xmin, xmax, ncells, name=""):
        e = np.zeros(max(int(ncells), 0) + 1)
        call("hrw_grid_log", float(xmin), float(xmax), int(ncells), _pd(e))
        self.scale = "log"
        self._compute(e, name)
        return self

    def geometric(self, xmin, xmax, ratio, ncells, name=""):
        e = np.zeros(max(int(ncells), 0) + 1)
        call("hrw_grid_geometric", float(xmin), float(xmax), float(ratio), int(ncells), _pd(e))
        self.scale = "geometric"
        self._compute(e, name)
        return self


# ---------------------------------------------------------------------------------- problems
class Problem:
    """A device problem descriptor (the user RHS of the reference examples)."""

    _h = None

    def __del__(self):
        _release(self, "hrw_problem_destroy")

    @property
    def neq(self) -> int:
        n = C.c_int64()
        call("hrw_problem_neq", self._h, C.byref(n))
        return n.value

    def tune(self, variant: int = 0, cells_per_lane: int = 0):
        """Kernel choice of the 1D FAST form (0: register-blocked, 1: LDS tile, also for the
        adjoint's reverse passes, 2: persistent; 0 with cells_per_lane 0: chosen by the line length);
        for a 2D problem, variant 0 and cells_per_lane = rows per wavefront (0: automatic)."""
        call("hrw_problem_tune", self._h, int(variant), int(cells_per_lane))
        return self

    def rhs(self, t, u, udot=None):
        torch = _torch()
        if udot is None:
            udot = torch.empty_like(u)
        n = self.neq
        call("hrw_problem_rhs", self._h, float(t), _dvec(u, n, "u"),
             _dvec(udot, n, "udot", u.device), _stream())
        return udot


class Burgers1D(Problem):
    """1D Burgers u_t + (u**2/2)_x = 0 (example/example1_burgers_1d_fv.f90:72-122).  With
    xedges (n+1 edges, zero-gradient boundaries) the reconstruction uses the non-uniform grid's
    own coefficients, as weno(ncells, k, eps, xedges) (src/hrweno_weno.f90:100-112)."""

    def __init__(self, n, k=3, eps=1e-6, flux=LAX_FRIEDRICHS, bc=BC_PERIODIC, alpha=-1.0,
                 xmin=-1.0, xmax=1.0, width=None, arith="fast", xedges=None):
        h = C.c_void_p()
        if xedges is not None:
            xe = _np(xedges)
            if xe.size != n + 1:
                raise HrwError(1, "Invalid input 'xedges': size(xedges) /= ncells + 1.")
            call("hrw_problem_burgers1d_edges", int(n), int(k), float(eps), int(flux), int(bc),
                 float(alpha), _pd(xe), _ARITH[arith], C.byref(h))
        else:
            w = _np(width) if width is not None else None
            call("hrw_problem_burgers1d", int(n), int(k), float(eps), int(flux), int(bc),
                 float(alpha), float(xmin), float(xmax), _pd(w) if w is not None else None,
                 _ARITH[arith], C.byref(h))
        self._h = h


class PBE2D(Problem):
    """2D population balance with growth rates g1 = g1a + g1b*x1, g2 = g2a + g2b*x2
    (example/example2_pbe_2d_fv.f90:73-155); state u((j-1)*n1 + i)."""

    def __init__(self, n1, n2, k=3, eps=1e-6, bc=BC_INFLOW0_OUTFLOW, x1=(0.0, 1.0),
                 x2=(0.0, 1.0), g1=(1.0, 0.5), g2=(0.5, 0.0), arith="fast"):
        h = C.c_void_p()
        call("hrw_problem_pbe2d", int(n1), int(n2), int(k), float(eps), int(bc), float(x1[0]),
             float(x1[1]), float(x2[0]), float(x2[1]), float(g1[0]), float(g1[1]),
             float(g2[0]), float(g2[1]), _ARITH[arith], C.byref(h))
        self._h = h


# ---------------------------------------------------------------------------------- ODEs
class _Ode:
    _h = None
    _cb = None

    def __del__(self):
        _release(self, "hrw_ode_destroy")

    def _wrap(self, fu):
        def trampoline(ctx, t, u, udot, stream):
            try:
                rc = fu(t, u, udot, stream)
                return int(rc or 0)
            except Exception:  # noqa: BLE001 -- reported as a non-zero status
                return 1
        self._cb = _lib.INTEGRAND(trampoline)
        return self._cb

    def _info(self):
        neq, order, fe, st = C.c_int64(), C.c_int(), C.c_int(), C.c_int()
        call("hrw_ode_info", self._h, C.byref(neq), C.byref(order), C.byref(fe), C.byref(st))
        return neq.value, order.value, fe.value, st.value

    neq = property(lambda self: self._info()[0])
    order = property(lambda self: self._info()[1])
    fevals = property(lambda self: self._info()[2])
    istate = property(lambda self: self._info()[3])

    def integrate(self, u, t: float, tout: float, dt: float, itask: int = 1) -> float:
        """``call ode%integrate(u, t, tout, dt[, itask])``: u (device tensor) is advanced in
        place; returns the new t (src/hrweno_tvdode.f90:97-178, :203-271)."""
        if u.numel() != self.neq or not u.is_contiguous():
            raise HrwError(1, "integrate: u must be a contiguous vector of neq values")
        tt = C.c_double(t)
        call("hrw_ode_set_stream", self._h, _stream())
        call("hrw_ode_integrate", self._h, _dvec(u, self.neq, "u"), C.byref(tt), float(tout),
             float(dt), int(itask))
        return tt.value

    def tape(self, ckpt_every: int = 1):
        """Record the following integrate calls for :meth:`adjoint` (0 stops).  rktvd records
        its last integrate call with a checkpoint every ``ckpt_every`` steps; mstvd keeps every
        state from the call that starts the method (istate 1) on, so tape must be set before
        that call.  Needs a device problem (1D Burgers on a linear grid)."""
        call("hrw_ode_tape", self._h, int(ckpt_every))
        return self

    def adjoint(self, lam):
        """Map lam = dJ/du at the end of the recording, in place, to dJ/du at its start:
        the start of the last integrate call (rktvd) or of the method (mstvd)."""
        call("hrw_ode_set_stream", self._h, _stream())
        call("hrw_ode_adjoint", self._h, _dvec(lam, self.neq, "lambda"))
        return lam

    def set_check(self, check_finite: bool = True):
        """Scan u for NaN/Inf at the end of every integrate call (the default); a hit raises
        HrwError and sets istate = -1."""
        call("hrw_ode_set_check", self._h, int(bool(check_finite)))
        return self

    def save(self, u, t: float) -> bytes:
        """The object's state (fevals, istate, the multistep history) with u and t, as bytes
        for resume in another object or process (:meth:`load`)."""
        nb = C.c_int64()
        call("hrw_ode_state_size", self._h, C.byref(nb))
        buf = C.create_string_buffer(nb.value)
        call("hrw_ode_set_stream", self._h, _stream())
        call("hrw_ode_save", self._h, _dvec(u, self.neq, "u"), float(t), buf, nb.value)
        return buf.raw

    def load(self, state: bytes, u) -> float:
        """Restore a state written by :meth:`save` (same method, order and neq) into this
        object and u; returns t."""
        tt = C.c_double()
        call("hrw_ode_set_stream", self._h, _stream())
        call("hrw_ode_load", self._h, _dvec(u, self.neq, "u"), C.byref(tt),
             C.c_char_p(bytes(state)), len(state))
        return tt.value


class RkTvd(_Ode):
    """``rktvd(fu, neq, order)`` (src/hrweno_tvdode.f90:69-95).

    ``fu`` is a Problem (fused device path) or a callable ``fu(t, u_ptr, udot_ptr, stream)``
    on raw device pointers (then ``neq`` is required)."""

    def __init__(self, fu, neq: Optional[int] = None, order: int = 3, arith="fast"):
        h = C.c_void_p()
        if isinstance(fu, Problem):
            call("hrw_rktvd_create", fu._h, int(order), C.byref(h))
            self._problem = fu
        else:
            call("hrw_rktvd_create_cb", self._wrap(fu), None, int(neq), int(order),
                 _ARITH[arith], C.byref(h))
        self._h = h


class MsTvd(_Ode):
    """``mstvd(fu, neq)`` (src/hrweno_tvdode.f90:180-201)."""

    def __init__(self, fu, neq: Optional[int] = None, arith="fast"):
        h = C.c_void_p()
        if isinstance(fu, Problem):
            call("hrw_mstvd_create", fu._h, C.byref(h))
            self._problem = fu
        else:
            call("hrw_mstvd_create_cb", self._wrap(fu), None, int(neq), _ARITH[arith],
                 C.byref(h))
        self._h = h


class BurgersBatch:
    """Ensemble of independent periodic 1D Burgers lines of uneven length, each with its own
    WENO order k, grid and time step, advanced together (one launch per RK3 stage for all
    lines, optionally from a captured hipGraph).  The state is one flat device tensor; line l
    is ``u[offsets[l]:offsets[l] + ncells[l]]``."""

    def __init__(self, ncells, k, dt, xmin=None, xmax=None, eps=1e-6, flux=LAX_FRIEDRICHS,
                 alpha=-1.0):
        n = np.ascontiguousarray(ncells, dtype=np.int32)
        kk = np.ascontiguousarray(k, dtype=np.int32)
        d = _np(dt)
        x0 = _np(xmin) if xmin is not None else None
        x1 = _np(xmax) if xmax is not None else None
        h = C.c_void_p()
        call("hrw_batch_burgers_create", n.size, n.ctypes.data_as(_lib.PI),
             kk.ctypes.data_as(_lib.PI), _pd(d), _pd(x0) if x0 is not None else None,
             _pd(x1) if x1 is not None else None, float(eps), int(flux), float(alpha), FAST,
             C.byref(h))
        self._h = h
        self.ncells, self.k, self.dt = n, kk, d
        total = C.c_int64()
        offs = np.zeros(n.size + 1, dtype=np.int64)
        call("hrw_batch_size", h, C.byref(total), offs.ctypes.data_as(_lib.PL))
        self.total, self.offsets = total.value, offs
        self.t = np.zeros(n.size)
        self._tstream = None

    # The library runs a batch on a non-default stream of its own (it captures hipGraphs).
    # Each call orders that stream after the caller's current stream and the caller's stream
    # after the batch's work, so torch ops and events on the caller's stream see the results.
    def _enter(self):
        torch = _torch()
        if self._tstream is None:
            self._tstream = torch.cuda.Stream()
            call("hrw_batch_set_stream", self._h, C.c_void_p(self._tstream.cuda_stream))
        cur = torch.cuda.current_stream()
        self._tstream.wait_stream(cur)
        return cur

    def _leave(self, cur):
        cur.wait_stream(self._tstream)

    def __del__(self):
        _release(self, "hrw_batch_destroy")

    def pack(self, lines, device):
        """Flat device state from a list of per-line host arrays."""
        torch = _torch()
        flat = np.zeros(self.total)
        for l, a in enumerate(lines):
            flat[self.offsets[l]:self.offsets[l] + self.ncells[l]] = a
        return torch.as_tensor(flat, device=device).clone()

    def unpack(self, u):
        flat = u.cpu().numpy()
        return [flat[self.offsets[l]:self.offsets[l] + self.ncells[l]].copy()
                for l in range(self.ncells.size)]

    def set_stream(self, stream):
        """Run the batch on a given torch.cuda.Stream (not the default stream)."""
        self._tstream = stream
        call("hrw_batch_set_stream", self._h, C.c_void_p(stream.cuda_stream))

    def _u(self, u, name="u"):
        p = _dvec(u, self.total, name)
        if u.data_ptr() % 16:
            raise HrwError(1, f"{name}: the batch state must be 16-byte aligned")
        return p

    def tune(self, u) -> int:
        r = C.c_int()
        pu = self._u(u)
        cur = self._enter()
        call("hrw_batch_tune", self._h, pu, C.byref(r))
        self._leave(cur)
        return r.value

    def set_cells_per_lane(self, r: int):
        call("hrw_batch_set_cells_per_lane", self._h, int(r))

    def set_form(self, form: int):
        """0: one wavefront per chunk; 1: persistent wavefronts walking the chunk table."""
        call("hrw_batch_set_form", self._h, int(form))

    def config(self):
        """(cells per lane, form) of the stage kernel."""
        r, f = C.c_int(), C.c_int()
        call("hrw_batch_config", self._h, C.byref(r), C.byref(f))
        return r.value, f.value

    def step(self, u, nsteps: int, graph: bool = True):
        pu = self._u(u)
        cur = self._enter()
        call("hrw_batch_step", self._h, pu, int(nsteps), _pd(self.t), int(bool(graph)))
        self._leave(cur)

    def step_per_line(self, u, nsteps: int):
        pu = self._u(u)
        cur = self._enter()
        call("hrw_batch_step_per_line", self._h, pu, int(nsteps), _pd(self.t))
        self._leave(cur)


class Adjoint:
    """Discrete adjoint of ``nsteps`` fixed rktvd steps of 1D Burgers on a linear grid: one
    line (a :class:`Burgers1D` problem with either flux and either kind of ends, any order,
    step ``dt``) or every line of a Lax-Friedrichs :class:`BurgersBatch` (RK3, each line's own
    dt).  ``forward`` records u0 -> u_N with a
    checkpoint every ``ckpt_every`` steps (0: the largest interval that fits in half of the free
    device memory, usually the whole run, so nothing is recomputed); ``backward`` maps dJ/du_N
    to dJ/du_0.  The reference has no adjoint (this is the added differentiation form of
    src/hrweno_tvdode.f90:136-171)."""

    def __init__(self, system, nsteps: int, dt: Optional[float] = None, order: int = 3,
                 ckpt_every: int = 0):
        h = C.c_void_p()
        self._system = system   # keeps the problem / batch alive
        self.batch = isinstance(system, BurgersBatch)
        if self.batch:
            call("hrw_batch_adjoint_create", system._h, int(nsteps), int(ckpt_every), C.byref(h))
        else:
            if dt is None:
                raise ValueError("dt is required for a single-line adjoint")
            call("hrw_adjoint_create", system._h, int(order), float(dt), int(nsteps),
                 int(ckpt_every), C.byref(h))
        self._h = h
        self._n = None
        self.token = 0          # bumps on every recorded forward

    def __del__(self):
        _release(self, "hrw_adjoint_destroy")

    def _sync_in(self):
        if self.batch:   # the batch runs on its own stream, ordered against the caller's
            self._cur = self._system._enter()
        else:
            call("hrw_adjoint_set_stream", self._h, _stream())

    def _sync_out(self):
        if self.batch:
            self._system._leave(self._cur)

    def _neq(self) -> int:
        if self._n is None:
            self._n = self.info()["neq"]
        return self._n

    def set_cells_per_lane(self, r: int):
        """Cells per lane of the reverse kernel (2 or 4); results agree to rounding."""
        call("hrw_adjoint_set_cells_per_lane", self._h, int(r))
        return self

    def info(self) -> dict:
        neq, ns, b = C.c_int64(), C.c_int64(), C.c_int64()
        ck = C.c_int()
        call("hrw_adjoint_info", self._h, C.byref(neq), C.byref(ns), C.byref(ck), C.byref(b))
        return {"neq": neq.value, "nsteps": ns.value, "ckpt_every": ck.value,
                "device_bytes": b.value}

    def forward(self, u0, uN=None):
        """u_N from u0 (device tensors), recording the checkpoints."""
        if uN is None:
            uN = _torch().empty_like(u0)
        n = self._neq()
        p0, pN = _dvec(u0, n, "u0"), _dvec(uN, n, "uN", u0.device)
        self._sync_in()
        call("hrw_adjoint_forward", self._h, p0, pN)
        self._sync_out()
        self.token += 1
        return uN

    def backward(self, gN, g0=None):
        """dJ/du_0 from dJ/du_N for the last recorded forward (g0 may be gN)."""
        if g0 is None:
            g0 = _torch().empty_like(gN)
        n = self._neq()
        pN, p0 = _dvec(gN, n, "gN"), _dvec(g0, n, "g0", gN.device)
        if g0.data_ptr() % 16:
            raise HrwError(1, "g0: the cotangent must be 16-byte aligned")
        self._sync_in()
        call("hrw_adjoint_backward", self._h, pN, p0)
        self._sync_out()
        return g0


def is_done(t: float, tout: float, dt: float) -> bool:
    """src/hrweno_tvdode.f90:273-284."""
    return (t - tout) * np.copysign(1.0, dt) > 0.0


IntegrandFn = Callable[[float, int, int, int], int]
