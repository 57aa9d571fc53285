"""Device discrete adjoint beyond periodic Lax-Friedrichs lines, through the C-ABI: the general
reverse pass (rhs_vjp_kernel, and rhs_vjp5_kernel for K = 3 without a line-wide alpha) for the
Godunov flux and zero-gradient ends under rktvd
(hrw_adjoint_*), and the recorded multistep integrator (hrw_ode_tape / hrw_ode_adjoint on an
mstvd object) for every composition -- against the oracle adjoint, the closed forms the oracle
is pinned by, and central differences of the device forward."""
import numpy as np
import pytest

from hrweno_hip import workloads as W

pytestmark = pytest.mark.gpu
# The reverse pass differentiates the map at the device forward's states, which the FAST
# arithmetic form reaches to ~1e-15 of the oracle's: agreement to 1e-10 of the largest
# component (as for the periodic Lax-Friedrichs reverse kernel).
TOL = 1e-10
LF, GOD = 0, 1
PER, ZG = 0, 1


def _case(n, seed, noise=1e-3):
    e, xc, w = W.linear_centres(-1.0, 1.0, n)
    u0 = W.smooth_step_ic(xc, W.Rng(seed)) + noise * np.random.default_rng(seed).standard_normal(n)
    dt = W.cfl_dt(u0, 2.0 / n)
    return u0, w, dt


def _dev(a, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev).clone()


def _close(g, ref, tol=TOL):
    return np.max(np.abs(g - ref)) <= tol * max(np.abs(ref).max(), 1e-300)


@pytest.mark.parametrize("flux,bc", [(LF, ZG), (GOD, PER), (GOD, ZG)])
@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("order", [1, 3])
def test_rk_adjoint_matches_oracle(oracle, hh, dev, flux, bc, k, order):
    for n in (1, 2, 3, 5, 64, 257, 1000, 4097):
        u0, w, dt = _case(n, 100 + n + k)
        steps = 4
        gN = np.random.default_rng(n).standard_normal(n)
        ou, og = oracle.burgers_rk_adjoint(u0, gN, w, nsteps=steps, dt=dt, k=k, order=order,
                                           flux=flux, bc=bc)
        adj = hh.Adjoint(hh.Burgers1D(n, k=k, flux=flux, bc=bc), nsteps=steps, dt=dt,
                         order=order, ckpt_every=3)
        uN = adj.forward(_dev(u0, dev)).cpu().numpy()
        g0 = adj.backward(_dev(gN, dev)).cpu().numpy()
        assert np.max(np.abs(uN - ou)) <= 1e-12 * max(np.abs(ou).max(), 1.0), n
        assert _close(g0, og), (n, np.max(np.abs(g0 - og)) / np.abs(og).max())


@pytest.mark.parametrize("alpha", [-1.0, 2.5])
def test_rk_adjoint_order2_zero_gradient(oracle, hh, dev, alpha):
    n, steps = 777, 6
    u0, w, dt = _case(n, 9)
    gN = np.cos(np.arange(n))
    _, og = oracle.burgers_rk_adjoint(u0, gN, w, nsteps=steps, dt=dt, order=2, alpha=alpha,
                                      bc=ZG)
    adj = hh.Adjoint(hh.Burgers1D(n, alpha=alpha, bc=ZG), nsteps=steps, dt=dt, order=2)
    adj.forward(_dev(u0, dev))
    assert _close(adj.backward(_dev(gN, dev)).cpu().numpy(), og)


def _ms_device(hh, dev, u0, gN, tout, dt, **kw):
    ode = hh.MsTvd(hh.Burgers1D(u0.size, **kw)).tape(1)
    u = _dev(u0, dev)
    t = ode.integrate(u, 0.0, tout, dt)
    lam = _dev(gN, dev)
    ode.adjoint(lam)
    return u.cpu().numpy(), lam.cpu().numpy(), t


@pytest.mark.parametrize("flux,bc", [(LF, PER), (LF, ZG), (GOD, PER), (GOD, ZG)])
@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("arith", ["exact", "fast"])
def test_ms_adjoint_matches_oracle(oracle, hh, dev, flux, bc, k, arith):
    for n in (1, 3, 64, 300, 4097):
        u0, w, dt = _case(n, 200 + n + k)
        gN = np.random.default_rng(n + 1).standard_normal(n)
        tout = 13.5 * dt
        ou, og, ot, N = oracle.burgers_ms_adjoint(u0, gN, w, tout=tout, dt=dt, k=k, flux=flux, bc=bc)
        assert N == 14
        u, g, t = _ms_device(hh, dev, u0, gN, tout, dt, k=k, flux=flux, bc=bc, arith=arith)
        assert t == ot
        if arith == "exact":
            assert np.array_equal(u, ou), n
        else:
            assert np.max(np.abs(u - ou)) <= 1e-12 * max(np.abs(ou).max(), 1.0), n
        assert _close(g, og), (n, np.max(np.abs(g - og)) / np.abs(og).max())


def test_ms_adjoint_fixed_alpha_and_calls_appended(oracle, hh, dev):
    """A recording spans every integrate call since the method started: two calls give the
    gradient of the whole run; only the start (four RK3 steps) when the horizon is inside it."""
    n = 500
    u0, w, dt = _case(n, 31)
    gN = np.sin(np.arange(n) * 0.3)
    ou, og, ot, N = oracle.burgers_ms_adjoint(u0, gN, w, tout=20.5 * dt, dt=dt, alpha=1.7, bc=ZG)
    ode = hh.MsTvd(hh.Burgers1D(n, alpha=1.7, bc=ZG)).tape(1)
    u = _dev(u0, dev)
    t = ode.integrate(u, 0.0, 9.5 * dt, dt)
    t = ode.integrate(u, t, 20.5 * dt, dt)
    assert t == ot
    lam = ode.adjoint(_dev(gN, dev)).cpu().numpy()
    assert _close(lam, og)
    _, og4, _, N4 = oracle.burgers_ms_adjoint(u0, gN, w, tout=0.5 * dt, dt=dt, alpha=1.7, bc=ZG)
    assert N4 == 4
    _, g4, _ = _ms_device(hh, dev, u0, gN, 0.5 * dt, dt, alpha=1.7, bc=ZG)
    assert _close(g4, og4)


@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("flux", [LF, GOD])
def test_conserved_total_and_fixed_ends_on_device(hh, dev, k, flux):
    """Periodic: the gradient of sum_i w_i u_N,i is w whatever the state (either integrator).
    Zero-gradient ends: the end cells never change, so the gradient of u_N at an end cell is
    zero everywhere else and 1 there (to the rounding of the integrator's coefficient sums)."""
    n = 3000
    u0, w, dt = _case(n, 41 + k, noise=0.05)
    adj = hh.Adjoint(hh.Burgers1D(n, k=k, flux=flux), nsteps=25, dt=dt)
    adj.forward(_dev(u0, dev))
    g = adj.backward(_dev(w, dev)).cpu().numpy()
    assert np.max(np.abs(g - w)) <= 1e-13 * np.abs(w).max()
    _, g, _ = _ms_device(hh, dev, u0, w, 20.5 * dt, dt, k=k, flux=flux)
    assert np.max(np.abs(g - w)) <= 1e-13 * np.abs(w).max()
    for end in (0, n - 1):
        e = np.zeros(n)
        e[end] = 1.0
        rest = np.arange(n) != end
        adj = hh.Adjoint(hh.Burgers1D(n, k=k, flux=flux, bc=ZG), nsteps=25, dt=dt)
        uN = adj.forward(_dev(u0, dev)).cpu().numpy()
        g = adj.backward(_dev(e, dev)).cpu().numpy()
        assert uN[end] == u0[end] and not g[rest].any() and abs(g[end] - 1.0) <= 4e-16
        uN, g, _ = _ms_device(hh, dev, u0, e, 20.5 * dt, dt, k=k, flux=flux, bc=ZG)
        assert uN[end] == u0[end] and not g[rest].any() and abs(g[end] - 1.0) <= 4e-16


@pytest.mark.parametrize("flux,bc,alpha", [(GOD, PER, -1.0), (GOD, ZG, -1.0), (LF, PER, 1.7),
                                           (LF, ZG, 1.7), (LF, PER, -1.0), (LF, ZG, -1.0)])
@pytest.mark.parametrize("k", [2, 3])
def test_blocked_reverse_matches_tile_reverse(hh, dev, flux, bc, alpha, k):
    """K = 2 and 3 take the register-blocked reverse pass (120 cells per wavefront; with alpha =
    max|s| its alpha-bar shares per block and the fixup launch); variant 1 of the problem keeps
    the LDS-tile pass.  With the EXACT forward (the
    same states under both) the gradients agree to rounding: rktvd (one cotangent and pointwise
    terms; periodic Lax-Friedrichs takes its own kernel there) and the multistep integrator
    (two cotangents), on lengths shorter than the stencil, around one and two chunks, around the
    first and last interior chunk, on grids with exact (2^m cells) and inexact edges.  The two
    passes round differently (reciprocals against divides, products of the weights against
    normalised weights), which the derivatives through the nonlinear weights amplify over the
    42 reverse stages to ~1e-12: held to TOL, as against the oracle."""
    for n in (1, 2, 4, 5, 123, 124, 125, 130, 248, 249, 251, 375, 1000, 1024, 20000, 1 << 20,
              (1 << 20) + 3):
        u0, w, dt = _case(n, 500 + n)
        gN = np.random.default_rng(n + 7).standard_normal(n)
        got = {}
        for variant in (0, 1):
            def prob():
                return hh.Burgers1D(n, k=k, flux=flux, bc=bc, alpha=alpha,
                                    arith="exact").tune(variant)
            r = []
            if not (flux == LF and bc == PER):
                adj = hh.Adjoint(prob(), nsteps=3, dt=dt, order=3)
                adj.forward(_dev(u0, dev))
                r.append(adj.backward(_dev(gN, dev)).cpu().numpy())
            ode = hh.MsTvd(prob()).tape(1)
            u = _dev(u0, dev)
            ode.integrate(u, 0.0, 13.5 * dt, dt)
            r.append(ode.adjoint(_dev(gN, dev)).cpu().numpy())
            got[variant] = r
        for g0, g1 in zip(got[0], got[1]):
            assert _close(g0, g1), (n, np.max(np.abs(g0 - g1)) / np.abs(g1).max())


def _richardson_fd(fwd, u0, target, j, h=1e-6):
    def d(hh_):
        e = np.zeros(u0.size)
        e[j] = hh_
        up, um = fwd(u0 + e), fwd(u0 - e)
        return np.sum((up - um) * ((up - target) + (um - target))) / (4 * hh_)
    return (4 * d(h) - d(2 * h)) / 3


@pytest.mark.parametrize("method", ["rk", "ms"])
def test_example1_composition_fd(hh, dev, method):
    """Example 1's own composition (Godunov, zero-gradient ends, WENO5) over 150 steps on 2000
    cells: the device gradient against Richardson central differences of the device forward,
    1e-6 relative per cell (components below 1 % of the largest held to 1e-8 of it: round-off
    of the differenced trajectories and the endpoint flux's kinks, src/hrweno_fluxes.f90:66-74).
    The multistep run amplifies perturbations about 1e9-fold over these steps (its gradient
    reaches ~1e9, and a change in the forward's last bits moves it several-fold), so its
    differenced trajectories carry ~steps * eps * max|g| / h, 1.5e-8 of the largest component,
    as rounding noise: its components below 10 % of the largest are held to 1e-7 of it."""
    n, steps = 2000, 150
    u0, w, dt = _case(n, 51)
    rng = np.random.default_rng(52)

    def fwd(u):
        ut = _dev(u, dev)
        if method == "rk":
            ode = hh.RkTvd(hh.Burgers1D(n, flux=GOD, bc=ZG), order=3)
        else:
            ode = hh.MsTvd(hh.Burgers1D(n, flux=GOD, bc=ZG))
        ode.integrate(ut, 0.0, (steps - 0.5) * dt, dt)
        return ut.cpu().numpy()

    target = fwd(u0 + 0.01 * rng.standard_normal(n))
    r = fwd(u0) - target
    if method == "rk":
        adj = hh.Adjoint(hh.Burgers1D(n, flux=GOD, bc=ZG), nsteps=steps, dt=dt)
        adj.forward(_dev(u0, dev))
        g = adj.backward(_dev(r, dev)).cpu().numpy()
    else:
        _, g, _ = _ms_device(hh, dev, u0, r, (steps - 0.5) * dt, dt, flux=GOD, bc=ZG)
    floor = (1e-2 if method == "rk" else 1e-1) * np.abs(g).max()
    cells = list(rng.choice(n, 24, replace=False)) + [0, 1, n - 2, n - 1]
    for j in cells:
        fd = _richardson_fd(fwd, u0, target, j)
        assert abs(g[j] - fd) <= 1e-6 * max(abs(fd), floor), (j, g[j], fd)


def test_tape_contracts(hh, dev):
    from hrweno_hip import HrwError
    n = 200
    u0, w, dt = _case(n, 61)
    # a multistep recording starts with the method
    ode = hh.MsTvd(hh.Burgers1D(n, bc=ZG))
    u = _dev(u0, dev)
    t = ode.integrate(u, 0.0, 6.5 * dt, dt)
    ode.tape(1)
    with pytest.raises(HrwError):
        ode.integrate(u, t, 9.5 * dt, dt)
    # loading a saved state drops the recording
    ode = hh.MsTvd(hh.Burgers1D(n, bc=ZG)).tape(1)
    u = _dev(u0, dev)
    t = ode.integrate(u, 0.0, 6.5 * dt, dt)
    state = ode.save(u, t)
    ode.adjoint(_dev(np.ones(n), dev))
    ode.load(state, u)
    with pytest.raises(HrwError):
        ode.adjoint(_dev(np.ones(n), dev))
    # no recording: an error, not a silent identity
    with pytest.raises(HrwError):
        hh.MsTvd(hh.Burgers1D(n)).adjoint(_dev(np.ones(n), dev))
    # non-uniform grids and callback integrands are refused
    e = np.linspace(-1.0, 1.0, n + 1) ** 3
    with pytest.raises(HrwError):
        hh.MsTvd(hh.Burgers1D(n, bc=ZG, xedges=e)).tape(1)
