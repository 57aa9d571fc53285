"""The oracle's discrete adjoint (oracle/hrweno_adjoint.c) pinned before it serves as the
checker of the device adjoint:

* the RHS vector-Jacobian product against central differences of the oracle RHS along random
  directions (every k, tiny and odd line lengths, alpha = max|v| and fixed);
* the whole reverse sweep against central differences of the REFERENCE's own modules (the
  compiled oracle/_ref build of src/hrweno_weno.f90 + src/hrweno_fluxes.f90 driving RK3), and of
  the oracle forward for orders 1 and 2;
* a closed form: periodic finite volumes conserve sum_i w_i u_i exactly, so for
  J = sum_i w_i u_N,i the gradient is w itself, whatever the state;
* each planted mistake (weights held fixed, a wrong sign of alpha, alpha held fixed) fails the
  finite-difference check.
"""
import numpy as np
import pytest

from hrweno_hip import workloads as W

H = 1e-6


def _line(n, seed, noise=0.1):
    e, xc, w = W.linear_centres(-1.0, 1.0, n)
    rng = np.random.default_rng(seed)
    v = W.smooth_step_ic(xc, W.Rng(seed)) + noise * rng.standard_normal(n)
    return v, w, rng


@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 257])
@pytest.mark.parametrize("alpha", [-1.0, 1.3])
def test_rhs_vjp_directional_fd(oracle, k, n, alpha):
    v, w, rng = _line(n, 10 * n + k)
    for _ in range(3):
        mu = rng.standard_normal(n)
        d = rng.standard_normal(n)
        g = oracle.burgers_rhs_vjp(v, mu, w, k=k, alpha=alpha)
        lp = oracle.burgers_rhs(v + H * d, w, k=k, alpha=alpha)
        lm = oracle.burgers_rhs(v - H * d, w, k=k, alpha=alpha)
        fd = mu @ (lp - lm) / (2 * H)
        scale = np.abs(mu).sum() * np.abs(oracle.burgers_rhs(v, w, k=k, alpha=alpha)).max() + abs(fd)
        assert abs(g @ d - fd) <= 1e-6 * max(abs(fd), 1e-6 * scale), (g @ d, fd)


def _loss_fd(fwd, u0, target, j):
    e = np.zeros(u0.size)
    e[j] = H
    up, um = fwd(u0 + e), fwd(u0 - e)
    # difference of squares without cancellation: (a+ - a-)(a+ + a-)/2
    return np.sum((up - um) * ((up - target) + (um - target))) / (4 * H)


def _setup(n, steps, seed):
    e, xc, w = W.linear_centres(-1.0, 1.0, n)
    u0 = W.smooth_step_ic(xc, W.Rng(seed))
    dt = W.cfl_dt(u0, 2.0 / n)
    rng = np.random.default_rng(seed)
    return u0, w, dt, rng


def test_adjoint_vs_reference_fd(oracle, ref):
    n, steps = 96, 30
    u0, w, dt, rng = _setup(n, steps, 5)
    fr = lambda u: ref.burgers_rk3(u, w, steps, dt)
    target = fr(u0 + 0.01 * rng.standard_normal(n))
    uN, g = oracle.burgers_rk_adjoint(u0, np.zeros(n), w, nsteps=steps, dt=dt)
    assert np.array_equal(uN, fr(u0))            # the same forward map, bit for bit
    _, g = oracle.burgers_rk_adjoint(u0, uN - target, w, nsteps=steps, dt=dt)
    for j in rng.choice(n, 16, replace=False):
        fd = _loss_fd(fr, u0, target, j)
        assert abs(g[j] - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(g).max()), (j, g[j], fd)


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("k", [2, 3])
def test_adjoint_vs_oracle_fd(oracle, order, k):
    n, steps = 50, 12
    u0, w, dt, rng = _setup(n, steps, 7 + order)
    fo = lambda u: oracle.burgers_rk_adjoint(u, np.zeros(n), w, nsteps=steps, dt=dt, k=k,
                                             order=order)[0]
    target = fo(u0 + 0.01 * rng.standard_normal(n))
    uN = fo(u0)
    _, g = oracle.burgers_rk_adjoint(u0, uN - target, w, nsteps=steps, dt=dt, k=k, order=order)
    for j in rng.choice(n, 10, replace=False):
        fd = _loss_fd(fo, u0, target, j)
        assert abs(g[j] - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(g).max()), (j, g[j], fd)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_adjoint_of_conserved_total_is_width(oracle, k):
    n, steps = 333, 25
    u0, w, dt, _ = _setup(n, steps, 11)
    _, g = oracle.burgers_rk_adjoint(u0, w, w, nsteps=steps, dt=dt, k=k)
    assert np.max(np.abs(g - w)) <= 1e-13 * np.abs(w).max()


@pytest.mark.parametrize("plant", [1, 2, 3])
def test_planted_mistakes_are_caught(oracle, ref, plant):
    n, steps = 96, 30
    u0, w, dt, rng = _setup(n, steps, 5)
    fr = lambda u: ref.burgers_rk3(u, w, steps, dt)
    target = fr(u0 + 0.01 * rng.standard_normal(n))
    oracle.adjoint_plant(plant)
    try:
        _, g = oracle.burgers_rk_adjoint(u0, fr(u0) - target, w, nsteps=steps, dt=dt)
    finally:
        oracle.adjoint_plant(0)
    worst = 0.0
    for j in range(n):   # alpha-bar lands near the argmax only: check every cell
        fd = _loss_fd(fr, u0, target, j)
        worst = max(worst, abs(g[j] - fd) / max(abs(fd), 1e-3 * np.abs(g).max()))
    assert worst > 1e-3
