"""Host-side contracts of the C-ABI on the device: argument checks before any launch, the
non-finite state check of integrate, serialised ODE state (resume in a fresh object, bit for
bit), and the batch's ordering against the NULL stream for C callers."""
import ctypes as C

import numpy as np
import pytest

from hrweno_hip import workloads as W

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev).clone()


def test_wrong_length_or_strided_tensors_rejected(hh, dev):
    import torch
    n = 300
    p = hh.Burgers1D(n)
    ode = hh.RkTvd(p)
    good = _t(np.ones(n), dev)
    short = _t(np.ones(n - 1), dev)
    strided = torch.ones(2 * n, dtype=torch.float64, device=dev)[::2]
    for bad in (short, strided):
        with pytest.raises(hh.HrwError):
            ode.integrate(bad, 0.0, 1e-3, 1e-3)
        with pytest.raises(hh.HrwError):
            p.rhs(0.0, bad)
    with pytest.raises(hh.HrwError):
        p.rhs(0.0, good, short)
    with pytest.raises(hh.HrwError):
        hh.lax_friedrichs(good, short)
    with pytest.raises(hh.HrwError):
        hh.godunov(good, good, out=short)
    w = hh.Weno(n)
    with pytest.raises(hh.HrwError):
        w.reconstruct(good, short)
    with pytest.raises(hh.HrwError):
        w.reconstruct_lines(good, 2, 1, n, good, good, 1, n)   # 2 lines need 2n values
    b = hh.BurgersBatch(np.array([100, 37]), np.array([3, 2]), np.array([1e-3, 1e-3]))
    with pytest.raises(hh.HrwError):
        b.step(_t(np.ones(b.total - 1), dev), 1)
    adj = hh.Adjoint(p, nsteps=2, dt=1e-3)
    with pytest.raises(hh.HrwError):
        adj.forward(short)
    # nothing above launched or changed state: the objects still work
    u = good.clone()
    ode.integrate(u, 0.0, 1e-3, 1e-3)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(u).all())


@pytest.mark.parametrize("method", ["rk", "ms"])
def test_nonfinite_state_is_an_error(hh, dev, method):
    """A NaN in the state ends integrate with HRW_ERR_STATE and istate = -1 (the reference's
    fatal-error state); with the check switched off the same call integrates on."""
    n = 4096
    e, xc, w = W.linear_centres(-1.0, 1.0, n)
    u0 = W.smooth_step_ic(xc, W.Rng(77))
    dt = W.cfl_dt(u0, 2.0 / n)
    u0[1234] = np.nan
    for problem in (hh.Burgers1D(n), hh.Burgers1D(n, flux=hh.GODUNOV, bc=hh.BC_ZERO_GRADIENT)):
        ode = hh.RkTvd(problem) if method == "rk" else hh.MsTvd(problem)
        u = _t(u0, dev)
        with pytest.raises(hh.HrwError, match="non-finite") as err:
            ode.integrate(u, 0.0, 10.5 * dt, dt)
        assert err.value.code == 4 and ode.istate == -1
        ode = (hh.RkTvd(problem) if method == "rk" else hh.MsTvd(problem)).set_check(False)
        u = _t(u0, dev)
        ode.integrate(u, 0.0, 10.5 * dt, dt)
        assert ode.istate == 2
    # a finite state passes the check
    ode = hh.RkTvd(hh.Burgers1D(n))
    u = _t(np.nan_to_num(u0), dev)
    ode.integrate(u, 0.0, 10.5 * dt, dt)
    assert ode.istate == 2


@pytest.mark.parametrize("kind", ["rk", "ms", "ms_pbe", "ms_callback"])
def test_saved_state_resumes_bitwise(hh, dev, kind):
    """An integration split by save -> fresh object -> load (the blob as plain bytes, as another
    process would read it) equals the unbroken run bit for bit, with the same t, fevals and
    istate; the multistep history travels in the blob."""
    import torch
    if kind == "ms_pbe":
        n1, n2 = 150, 90
        e1, c1, _ = W.linear_centres(0.0, 1.0, n1)
        e2, c2, _ = W.linear_centres(0.0, 1.0, n2)
        X1, X2 = np.meshgrid(c1, c2, indexing="xy")
        u0 = np.exp(-((X1 - 0.3) ** 2 + (X2 - 0.35) ** 2) / (2 * 0.05 ** 2)).ravel()
        dt = 0.4 / (1.5 * n1 + 0.5 * n2)
        make = lambda: hh.MsTvd(hh.PBE2D(n1, n2))   # noqa: E731
    else:
        n = 3000
        e, xc, w = W.linear_centres(-1.0, 1.0, n)
        u0 = W.smooth_step_ic(xc, W.Rng(5))
        dt = W.cfl_dt(u0, 2.0 / n)
        if kind == "rk":
            make = lambda: hh.RkTvd(hh.Burgers1D(n, arith="exact"))   # noqa: E731
        elif kind == "ms":
            make = lambda: hh.MsTvd(hh.Burgers1D(n, arith="exact"))   # noqa: E731
        else:
            prob = hh.Burgers1D(n, arith="exact")

            def fu(t, up, udp, stream):
                return hh._lib.lib().hrw_problem_rhs(prob._h, t, up, udp, stream)
            make = lambda: hh.MsTvd(fu, neq=n, arith="exact")   # noqa: E731
    t1 = W.steps_tout(0.0, dt, 9)
    t2 = W.steps_tout(0.0, dt, 17)
    whole = make()
    uw = _t(u0, dev)
    tw = whole.integrate(uw, 0.0, t1, dt)
    tw = whole.integrate(uw, tw, t2, dt)
    first = make()
    ua = _t(u0, dev)
    ta = first.integrate(ua, 0.0, t1, dt)
    blob = bytes(first.save(ua, ta))
    del first
    second = make()
    ub = _t(np.zeros_like(u0), dev)
    tb = second.load(blob, ub)
    assert tb == ta
    tb = second.integrate(ub, tb, t2, dt)
    torch.cuda.synchronize()
    assert tb == tw and second.fevals == whole.fevals and second.istate == whole.istate
    assert torch.equal(ub, uw)


def test_saved_state_of_another_ode_is_refused(hh, dev):
    a = hh.MsTvd(hh.Burgers1D(100))
    u = _t(np.ones(100), dev)
    blob = a.save(u, 0.0)
    with pytest.raises(hh.HrwError):
        hh.RkTvd(hh.Burgers1D(100)).load(blob, u)
    with pytest.raises(hh.HrwError):
        hh.MsTvd(hh.Burgers1D(101)).load(blob, _t(np.ones(101), dev))
    with pytest.raises(hh.HrwError):
        a.load(b"x" * len(blob), u)


def test_batch_orders_with_null_stream_for_c_callers(hh, dev):
    """C-level sequence with no stream set: upload on the NULL stream, hrw_batch_step, then a
    NULL-stream copy back -- the copy sees the finished step (compared with the per-line path,
    bit for bit)."""
    lib = hh._lib.lib()
    rng = np.random.default_rng(4)
    lens = np.rint(np.exp(rng.uniform(np.log(256), np.log(40000), 64))).astype(np.int32)
    ks = np.where(rng.uniform(size=64) < 0.5, 2, 3).astype(np.int32)
    u0s = [W.smooth_step_ic(W.linear_centres(-1.0, 1.0, int(n))[1], W.Rng(900 + l))
           for l, n in enumerate(lens)]
    dts = np.array([W.cfl_dt(u, 2.0 / u.size) for u in u0s])
    outs = []
    for per_line in (False, True):
        h = C.c_void_p()
        assert lib.hrw_batch_burgers_create(64, lens.ctypes.data_as(hh._lib.PI),
                                            ks.ctypes.data_as(hh._lib.PI),
                                            dts.ctypes.data_as(hh._lib.PD), None, None, 1e-6, 0,
                                            -1.0, 1, C.byref(h)) == 0
        total = C.c_int64()
        offs = np.zeros(65, dtype=np.int64)
        lib.hrw_batch_size(h, C.byref(total), offs.ctypes.data_as(hh._lib.PL))
        flat = np.zeros(total.value)
        for l, u in enumerate(u0s):
            flat[offs[l]:offs[l] + u.size] = u
        d = C.c_void_p()
        assert lib.hrw_device_alloc(flat.nbytes, C.byref(d)) == 0
        assert lib.hrw_memcpy_htod(d, flat.ctypes.data_as(C.c_void_p), flat.nbytes, None) == 0
        rc = (lib.hrw_batch_step_per_line(h, d, 6, None) if per_line
              else lib.hrw_batch_step(h, d, 6, None, 1))
        assert rc == 0
        back = np.empty_like(flat)
        assert lib.hrw_memcpy_dtoh(back.ctypes.data_as(C.c_void_p), d, flat.nbytes, None) == 0
        outs.append(back)
        assert lib.hrw_batch_synchronize(h) == 0
        lib.hrw_device_free(d)
        lib.hrw_batch_destroy(h)
    assert np.array_equal(outs[0], outs[1])
    assert not np.array_equal(outs[0][offs[0]:offs[0] + lens[0]], u0s[0])
