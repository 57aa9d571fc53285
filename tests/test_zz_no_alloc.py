"""hrw_rktvd_steps_noalloc (stage buffers and alpha slots in a static device array) against
hrw_ode_integrate on hrw_rktvd_create of the same problem, through the C-ABI: bit for bit, two
calls in a row (the second on a smaller line, so stale scratch contents from the first would
show), every RK order and both fluxes; problems beyond the array's size are refused."""
import ctypes as C

import numpy as np
import pytest

from hrweno_hip import workloads as W

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev).clone()


def _steps(hh, dev, n, order, flux, bc, nsteps, seed):
    import torch
    lib = hh._lib.lib()
    e, xc, w = W.linear_centres(-1.0, 1.0, n)
    u0 = W.smooth_step_ic(xc, W.Rng(seed))
    dt = W.cfl_dt(u0, 2.0 / n)
    prob = hh.Burgers1D(n, 3, 1e-6, flux=flux, bc=bc)
    ode = hh.RkTvd(prob, order=order)
    ref = _t(u0, dev)
    t = ode.integrate(ref, 0.0, W.steps_tout(0.0, dt, nsteps), dt)
    assert ode.fevals == order * nsteps
    got = _t(u0, dev)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.hrw_rktvd_steps_noalloc(prob._h, order, C.c_void_p(got.data_ptr()), nsteps, dt, stream)
    assert rc == 0, lib.hrw_last_error().decode()
    torch.cuda.synchronize()
    return got, ref


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("flux,bc", [(0, 0), (1, 1)])
def test_noalloc_steps_equal_integrate(hh, dev, order, flux, bc):
    import torch
    for n, nsteps, seed in ((1 << 20, 7, 3), (1000, 11, 4)):   # the second call is smaller
        got, ref = _steps(hh, dev, n, order, flux, bc, nsteps, seed)
        assert torch.equal(got, ref), (n, order, flux)


def test_noalloc_refuses_larger_problems(hh, dev):
    import torch
    lib = hh._lib.lib()
    n = (1 << 20) + 1
    prob = hh.Burgers1D(n)
    u = _t(np.ones(n), dev)
    rc = lib.hrw_rktvd_steps_noalloc(prob._h, 3, C.c_void_p(u.data_ptr()), 1, 1e-6, None)
    assert rc == 1 and "exceed" in lib.hrw_last_error().decode()
    torch.cuda.synchronize()
    assert bool((u == 1.0).all())
