"""The oracle's discrete adjoint beyond periodic Lax-Friedrichs (oracle/hrweno_adjoint.c): the
Godunov flux, zero-gradient ends and the multistep integrator with its four RK3 starting steps,
pinned before they serve as the checker of the device adjoint:

* the forward maps the reverse sweeps differentiate are the reference's own: example 1's
  composition on the compiled reference modules (oracle/_ref: src/hrweno_weno.f90 and
  src/hrweno_fluxes.f90, the integrator lines restated) gives the oracle's states bit for bit;
* the RHS vector-Jacobian product against central differences of the oracle RHS along random
  directions, for both fluxes and both kinds of ends;
* the whole reverse sweep (rktvd and mstvd) against central differences of the reference build;
* closed forms: periodic finite volumes conserve sum_i w_i u_i under either flux and either
  integrator, so the gradient of that total is w; with zero-gradient ends the reference copies
  the boundary fluxes inward (example/example1_burgers_1d_fv.f90:103-104), so the two end cells
  never change and the gradient of u_N at an end cell is the unit vector there;
* each planted mistake (Godunov differentiated on the other end state, the end cells' RHS kept,
  end-cell cotangents wrapped around the line, the multistep history state or slope dropped)
  fails one of these checks.
"""
import numpy as np
import pytest

from hrweno_hip import workloads as W

H = 1e-6
LF, GOD = 0, 1
PER, ZG = 0, 1


def _setup(n, seed, noise=1e-3):
    """A smooth step with a little noise: the step alone is mirror-symmetric, so |u| would tie
    at its maximum and Lax-Friedrichs alpha = max|u| would have a kink on the trajectory."""
    e, xc, w = W.linear_centres(-1.0, 1.0, n)
    u0 = W.smooth_step_ic(xc, W.Rng(seed))
    rng = np.random.default_rng(seed)
    if noise:
        u0 = u0 + noise * rng.standard_normal(n)
    dt = W.cfl_dt(u0, 2.0 / n)
    return u0, w, dt, rng


def _loss_fd(fwd, u0, target, j):
    """d/du0_j of |u_N - target|^2 / 2 by central differences, Richardson-extrapolated:
    (4 D(h) - D(2h)) / 3 cancels the h^2 term, which the sharply turning weights (eps = 1e-6)
    make visible at 1e-6 relative."""
    def d(h):
        e = np.zeros(u0.size)
        e[j] = h
        up, um = fwd(u0 + e), fwd(u0 - e)
        return np.sum((up - um) * ((up - target) + (um - target))) / (4 * h)
    return (4 * d(H) - d(2 * H)) / 3


def _moves_argmax(u0, j):
    """Whether a +-2h perturbation of cell j moves the first argmax of |u|: Lax-Friedrichs
    alpha = max|u| has a kink there (a tie at the maximum), where central differences average
    two one-sided slopes and the adjoint takes the first cell's."""
    e = np.zeros(u0.size)
    e[j] = 2 * H
    return len({int(np.argmax(np.abs(u0 + s * e))) for s in (-1.0, 0.0, 1.0)}) > 1


def _check_fd(fwd, g, u0, target, cells, flux, rel=1e-6):
    # Components below 1 % of the largest are held to 1e-8 of the largest: the differenced
    # trajectories (tens of steps) carry ~1e-10 of round-off at h = 1e-6, and Godunov's endpoint
    # flux has kinks where |vm| = |vp| (src/hrweno_fluxes.f90:66-74), which some interface
    # passes within +-2h over that many steps
    floor = 1e-2 * np.abs(g).max()
    checked = 0
    for j in cells:
        if flux == LF and _moves_argmax(u0, j):
            continue
        fd = _loss_fd(fwd, u0, target, j)
        assert abs(g[j] - fd) <= rel * max(abs(fd), floor), (j, g[j], fd)
        checked += 1
    assert checked >= len(cells) - 2


@pytest.mark.parametrize("flux,bc", [(LF, ZG), (GOD, PER), (GOD, ZG)])
@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 257])
def test_rhs_vjp_directional_fd(oracle, flux, bc, k, n):
    e, xc, w = W.linear_centres(-1.0, 1.0, n)
    rng = np.random.default_rng(10 * n + k + 100 * flux)
    v = W.smooth_step_ic(xc, W.Rng(n + k)) + 0.1 * rng.standard_normal(n)
    kw = dict(k=k, flux=flux, bc=bc)
    for _ in range(3):
        mu = rng.standard_normal(n)
        d = rng.standard_normal(n)
        g = oracle.burgers_rhs_vjp(v, mu, w, **kw)
        dd = lambda h: mu @ (oracle.burgers_rhs(v + h * d, w, **kw) -
                             oracle.burgers_rhs(v - h * d, w, **kw)) / (2 * h)
        fd = (4 * dd(H) - dd(2 * H)) / 3   # Richardson: the h^2 term cancels
        scale = np.abs(mu).sum() * np.abs(oracle.burgers_rhs(v, w, **kw)).max() + abs(fd)
        assert abs(g @ d - fd) <= 1e-6 * max(abs(fd), 1e-6 * scale), (g @ d, fd)


@pytest.mark.parametrize("flux,bc", [(LF, PER), (LF, ZG), (GOD, PER), (GOD, ZG)])
def test_forward_maps_are_the_reference(oracle, ref, flux, bc):
    """The states both reverse sweeps differentiate, against example 1's composition on the
    reference's own modules, bit for bit."""
    n = 80
    u0, w, dt, _ = _setup(n, 3)
    steps = 9
    ou, _ = oracle.burgers_rk_adjoint(u0, np.zeros(n), w, nsteps=steps, dt=dt, flux=flux, bc=bc)
    assert np.array_equal(ou, ref.burgers_ex(u0, w, steps, dt, flux=flux, bc=bc, method=3))
    # the oracle's own integrator driver takes the same steps
    pu, _, psteps, _ = oracle.burgers_rk(u0, w, flux=flux, bc=bc, tout=(steps - 0.5) * dt, dt=dt)
    assert psteps == steps and np.array_equal(pu, ou)
    mu, mg, t, N = oracle.burgers_ms_adjoint(u0, np.zeros(n), w, tout=11.5 * dt, dt=dt, flux=flux,
                                             bc=bc)
    assert N == 12
    assert np.array_equal(mu, ref.burgers_ex(u0, w, N, dt, flux=flux, bc=bc, method=0))
    su, st, _ = oracle.burgers_ms(u0, w, flux=flux, bc=bc, tout=11.5 * dt, dt=dt)
    assert st == t and np.array_equal(su, mu)


@pytest.mark.parametrize("flux,bc", [(LF, ZG), (GOD, PER), (GOD, ZG)])
def test_rk_adjoint_vs_reference_fd(oracle, ref, flux, bc):
    n, steps = 96, 30
    u0, w, dt, rng = _setup(n, 5)
    fr = lambda u: ref.burgers_ex(u, w, steps, dt, flux=flux, bc=bc, method=3)
    target = fr(u0 + 0.01 * rng.standard_normal(n))
    uN, g = oracle.burgers_rk_adjoint(u0, np.zeros(n), w, nsteps=steps, dt=dt, flux=flux, bc=bc)
    _, g = oracle.burgers_rk_adjoint(u0, uN - target, w, nsteps=steps, dt=dt, flux=flux, bc=bc)
    _check_fd(fr, g, u0, target, list(rng.choice(n, 14, replace=False)) + [0, n - 1], flux)


@pytest.mark.parametrize("flux,bc", [(LF, PER), (LF, ZG), (GOD, PER), (GOD, ZG)])
def test_ms_adjoint_vs_reference_fd(oracle, ref, flux, bc):
    n = 96
    u0, w, dt, rng = _setup(n, 6)
    tout = 29.5 * dt
    _, _, _, N = oracle.burgers_ms_adjoint(u0, np.zeros(n), w, tout=tout, dt=dt, flux=flux, bc=bc)
    assert N == 30
    fr = lambda u: ref.burgers_ex(u, w, N, dt, flux=flux, bc=bc, method=0)
    target = fr(u0 + 0.01 * rng.standard_normal(n))
    uN = fr(u0)
    _, g, _, _ = oracle.burgers_ms_adjoint(u0, uN - target, w, tout=tout, dt=dt, flux=flux, bc=bc)
    _check_fd(fr, g, u0, target, list(rng.choice(n, 14, replace=False)) + [0, n - 1], flux)


@pytest.mark.parametrize("k", [1, 2])
@pytest.mark.parametrize("flux,bc", [(LF, PER), (GOD, ZG)])
def test_ms_adjoint_vs_oracle_fd(oracle, k, flux, bc):
    """Orders k = 1, 2 (the reference build runs k = 3 above) against the oracle's own mstvd."""
    n = 50
    u0, w, dt, rng = _setup(n, 8 + k)
    tout = 17.5 * dt
    fo = lambda u: oracle.burgers_ms(u, w, k=k, flux=flux, bc=bc, tout=tout, dt=dt)[0]
    target = fo(u0 + 0.01 * rng.standard_normal(n))
    _, g, _, N = oracle.burgers_ms_adjoint(u0, fo(u0) - target, w, tout=tout, dt=dt, k=k,
                                           flux=flux, bc=bc)
    assert N == 18
    _check_fd(fo, g, u0, target, rng.choice(n, 10, replace=False), flux)


@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("flux", [LF, GOD])
def test_conserved_total_gradient_is_width(oracle, k, flux):
    n = 333
    u0, w, dt, _ = _setup(n, 11, noise=0.05)
    _, g = oracle.burgers_rk_adjoint(u0, w, w, nsteps=25, dt=dt, k=k, flux=flux)
    assert np.max(np.abs(g - w)) <= 1e-13 * np.abs(w).max()
    _, g, _, N = oracle.burgers_ms_adjoint(u0, w, w, tout=20.5 * dt, dt=dt, k=k, flux=flux)
    assert N == 21
    assert np.max(np.abs(g - w)) <= 1e-13 * np.abs(w).max()


@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("flux", [LF, GOD])
def test_zero_gradient_end_cells_are_fixed(oracle, k, flux):
    n = 120
    u0, w, dt, _ = _setup(n, 12, noise=0.05)
    # (exactly zero away from the end cell; the end cell's 1 comes through the integrator's
    # pointwise coefficients, whose sum is 1 to rounding)
    for end in (0, n - 1):
        e = np.zeros(n)
        e[end] = 1.0
        rest = np.arange(n) != end
        uN, g = oracle.burgers_rk_adjoint(u0, e, w, nsteps=20, dt=dt, k=k, flux=flux, bc=ZG)
        assert uN[end] == u0[end] and not g[rest].any() and abs(g[end] - 1.0) <= 4e-16
        uN, g, _, _ = oracle.burgers_ms_adjoint(u0, e, w, tout=20.5 * dt, dt=dt, k=k, flux=flux,
                                                bc=ZG)
        assert uN[end] == u0[end] and not g[rest].any() and abs(g[end] - 1.0) <= 4e-16


def test_ms_adjoint_without_steps_and_short_horizons(oracle):
    n = 40
    u0, w, dt, rng = _setup(n, 13)
    gN = rng.standard_normal(n)
    # is_done at the start: no step, nothing changes (src/hrweno_tvdode.f90:229)
    u, g, t, N = oracle.burgers_ms_adjoint(u0, gN, w, t0=1.0, tout=0.5, dt=dt)
    assert N == 0 and t == 1.0 and np.array_equal(u, u0) and np.array_equal(g, gN)
    # a horizon inside the start: the four RK3 starting steps are taken all the same
    u, g, t, N = oracle.burgers_ms_adjoint(u0, gN, w, tout=0.5 * dt, dt=dt)
    assert N == 4
    ur, gr = oracle.burgers_rk_adjoint(u0, gN, w, nsteps=4, dt=dt)
    assert np.array_equal(u, ur)
    # no history term reaches back past the start: the reverse is the RK3 one
    assert np.max(np.abs(g - gr)) <= 1e-14 * np.abs(gr).max()


def _worst_ref_fd(fr, grad, u0, target, n, flux):
    """The largest disagreement over every cell, by the rules of _check_fd."""
    worst = 0.0
    g = grad()
    floor = 1e-2 * np.abs(g).max()
    for j in range(n):
        if flux == LF and _moves_argmax(u0, j):
            continue
        fd = _loss_fd(fr, u0, target, j)
        worst = max(worst, abs(g[j] - fd) / max(abs(fd), floor))
    return worst


@pytest.mark.parametrize("plant,method,flux,bc", [(4, 3, GOD, ZG), (5, 3, GOD, ZG), (6, 3, LF, ZG),
                                                  (7, 0, LF, PER), (8, 0, GOD, ZG)])
def test_planted_mistakes_are_caught(oracle, ref, plant, method, flux, bc):
    n = 64
    u0, w, dt, rng = _setup(n, 5)
    steps = 24
    fr = lambda u: ref.burgers_ex(u, w, steps, dt, flux=flux, bc=bc, method=method)
    target = fr(u0 + 0.01 * rng.standard_normal(n))
    r = fr(u0) - target
    if method == 3:
        grad = lambda: oracle.burgers_rk_adjoint(u0, r, w, nsteps=steps, dt=dt, flux=flux, bc=bc)[1]
    else:
        grad = lambda: oracle.burgers_ms_adjoint(u0, r, w, tout=(steps - 0.5) * dt, dt=dt,
                                                 flux=flux, bc=bc)[1]
    # the exact adjoint over every cell (Godunov's kinks can reach 1e-6 here), against > 1e-3
    assert _worst_ref_fd(fr, grad, u0, target, n, flux) < 1e-5
    oracle.adjoint_plant(plant)
    try:
        worst = _worst_ref_fd(fr, grad, u0, target, n, flux)
    finally:
        oracle.adjoint_plant(0)
    assert worst > 1e-3, worst
