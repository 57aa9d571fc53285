"""Fused device integrators (one pass per stage) vs the oracle, through the C-ABI.

EXACT arithmetic: bit-identical to the oracle (numeric equality of every cell, same t, same
step count and fevals).  FAST arithmetic: per-cell difference within 1e-12 of max|u|, and
within 1e-12 relative on every cell with |u| >= 1e-3 max|u| (the tolerance the fast form is
held to; it differs from EXACT by FMA reordering and the common-denominator weights).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _t(a, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev).clone()


def _h(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def assert_close(g, c, tol=TOL):
    scale = np.abs(c).max()
    d = np.abs(g - c)
    assert d.max() <= tol * scale, d.max() / scale
    big = np.abs(c) >= 1e-3 * scale
    assert np.max(d[big] / np.abs(c[big])) <= tol


def run_burgers(hh, dev, u0, *, n, k=3, flux=0, bc=0, order=3, arith="exact", t0=0.0,
                tout=0.5, dt=1e-3, itask=1, method="rk", xmin=-1.0, xmax=1.0):
    p = hh.Burgers1D(n, k, 1e-6, flux=flux, bc=bc, xmin=xmin, xmax=xmax, arith=arith)
    ode = hh.RkTvd(p, order=order) if method == "rk" else hh.MsTvd(p)
    u = _t(u0, dev)
    t = ode.integrate(u, t0, tout, dt, itask)
    return _h(u), t, ode


def test_config1_exact_bitwise(oracle, hh, dev):
    """1000 cells, WENO5 + LF(alpha=max|u|) + RK3, periodic, CFL 0.4, to t = 0.5."""
    from hrweno_hip import workloads as W
    c = W.config1()
    u, t, ode = run_burgers(hh, dev, c.u0, n=c.n, tout=c.tout, dt=c.dt)
    ou, ot, steps, fe = oracle.burgers_rk(c.u0, c.width, tout=c.tout, dt=c.dt)
    assert steps > 900 and t == ot and ode.fevals == fe and ode.istate == 2
    assert np.array_equal(u, ou)


def test_config1_fast_within_tolerance(oracle, hh, dev):
    from hrweno_hip import workloads as W
    c = W.config1()
    u, t, ode = run_burgers(hh, dev, c.u0, n=c.n, tout=c.tout, dt=c.dt, arith="fast")
    ou, ot, steps, fe = oracle.burgers_rk(c.u0, c.width, tout=c.tout, dt=c.dt)
    assert t == ot and ode.fevals == fe
    assert_close(u, ou)


@pytest.mark.parametrize("n", [1, 3, 7, 1021, 1022, 1023, 2044, 2045, 5000])
def test_tiling_invariance_exact(oracle, hh, dev, n):
    """Tile edges (1022 cells per workgroup) and tiny lines change nothing."""
    rng = np.random.default_rng(n)
    u0 = 0.5 + rng.standard_normal(n) * 0.3
    e = -1.0 + (2.0 / n) * np.arange(n + 1)
    w = e[1:] - e[:-1]
    dt = 0.2 * (2.0 / n) / np.abs(u0).max()
    u, t, ode = run_burgers(hh, dev, u0, n=n, tout=7.5 * dt, dt=dt)
    ou, ot, steps, fe = oracle.burgers_rk(u0, w, tout=7.5 * dt, dt=dt)
    assert steps == 8 and t == ot
    assert np.array_equal(u, ou)


@pytest.mark.parametrize("bc", [0, 1])
@pytest.mark.parametrize("k", [1, 2, 3])
def test_fast_kernels_all_sizes(oracle, hh, dev, k, bc):
    """FAST form through the 1D kernels (register-blocked with 4, 6 or 8 cells per lane or
    chosen by the line length, its persistent prefetching form, and the LDS tile) on line
    lengths that
    hit every edge/interior case (chunks of 62R output cells with R + K - 1 halo cells before
    and 63R + K - 1 after): each within the FAST tolerance of the oracle, and the
    register-blocked forms bit-identical (the per-cell arithmetic does not depend on which
    lane computes it)."""
    for n in (1, 2, 5, 37, 247, 248, 255, 256, 378, 507, 1000, 2049, 2050, 4093, 20000, 300001):
        rng = np.random.default_rng(1000 * k + n)
        u0 = 0.4 + rng.standard_normal(n) * 0.3 + (np.arange(n) > n // 3)
        e = -1.0 + (2.0 / n) * np.arange(n + 1)
        w = e[1:] - e[:-1]
        dt = 0.3 * (2.0 / n) / np.abs(u0).max()
        ou, ot, steps, fe = oracle.burgers_rk(u0, w, k=k, bc=bc, flux=0 if bc == 0 else 1,
                                              tout=5.5 * dt, dt=dt)
        res = {}
        for variant, r in ((0, 8), (0, 4), (0, 6), (0, 0), (2, 4), (2, 6), (2, 8), (1, 8)):
            p = hh.Burgers1D(n, k, 1e-6, flux=0 if bc == 0 else 1, bc=bc, arith="fast")
            p.tune(variant, r)
            ode = hh.RkTvd(p, order=3)
            u = _t(u0, dev)
            t = ode.integrate(u, 0.0, 5.5 * dt, dt)
            assert t == ot and ode.fevals == fe
            res[(variant, r)] = _h(u)
            assert_close(res[(variant, r)], ou)
        for key in ((0, 4), (0, 6), (0, 0), (2, 4), (2, 6), (2, 8)):
            assert np.array_equal(res[(0, 8)], res[key]), (n, key)


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("k", [1, 2, 3])
def test_orders_and_k_exact(oracle, hh, dev, order, k):
    from hrweno_hip import workloads as W
    c = W.config1(n=600, seed=21)
    u, t, ode = run_burgers(hh, dev, c.u0, n=c.n, k=k, order=order, tout=0.05, dt=c.dt)
    ou, ot, steps, fe = oracle.burgers_rk(c.u0, c.width, k=k, order=order, tout=0.05, dt=c.dt)
    assert t == ot and ode.fevals == fe
    assert np.array_equal(u, ou)


def test_grid_far_from_origin(oracle, hh, dev):
    """A linear grid whose edges lie 1e12 cell widths from the origin (widths known to ~1e-4
    only): the FAST form leaves the register kernels, whose 1/width starts from 1/rx, for the
    tile kernel, and agrees with the oracle on the same widths."""
    n = 1000
    xmin, xmax = 1e9, 1e9 + 1.0
    e = xmin + ((xmax - xmin) / n) * np.arange(n + 1)
    w = e[1:] - e[:-1]
    xc = np.linspace(-1.0, 1.0, n)
    u0 = 0.5 + np.sin(np.pi * xc)
    dt = 0.2 * w.min() / np.abs(u0).max()
    ou, ot, steps, fe = oracle.burgers_rk(u0, w, tout=30 * dt, dt=dt)
    for arith in ("exact", "fast"):
        u, t, ode = run_burgers(hh, dev, u0, n=n, tout=30 * dt, dt=dt, xmin=xmin, xmax=xmax,
                                arith=arith)
        assert t == ot and ode.fevals == fe
        if arith == "exact":
            assert np.array_equal(u, ou)
        else:
            assert_close(u, ou)


@pytest.mark.parametrize("flux", [0, 1])
@pytest.mark.parametrize("k", [1, 2, 3])
def test_nonuniform_grid_fused(oracle, hh, dev, k, flux):
    """weno(ncells, k, eps, xedges) in the example-1 composition (zero-gradient boundary
    fluxes) on a smoothly stretched grid: the fused problem built from the edges equals the
    oracle run with the grid's cnu bit for bit in the EXACT form and within the FAST
    tolerance, for rktvd and mstvd; on this grid the uniform coefficients give another answer,
    so the test sees the grid's own."""
    n = 300
    s = np.linspace(0.0, 1.0, n + 1)
    xe = -5.0 + 10.0 * (s + 0.6 * np.sin(2 * np.pi * s) / (2 * np.pi))   # increasing, stretched
    w = xe[1:] - xe[:-1]
    xc = (xe[1:] + xe[:-1]) / 2
    u0 = np.clip(1.0 - 0.25 * (xc + 4.0), -0.5, 1.0) + 0.1 * np.sin(xc)
    dt = 0.3 * w.min() / np.abs(u0).max()
    cnu = oracle.calc_cnu(k, xe)
    ou, ot, steps, fe = oracle.burgers_rk(u0, w, k=k, flux=flux, bc=1, tout=40 * dt, dt=dt,
                                          cnu=cnu)
    if k > 1:
        uu, *_ = oracle.burgers_rk(u0, w, k=k, flux=flux, bc=1, tout=40 * dt, dt=dt)
        assert np.abs(uu - ou).max() > 1e-6
    om, omt, ofe = oracle.burgers_ms(u0, w, k=k, flux=flux, bc=1, tout=30 * dt, dt=dt, cnu=cnu)
    for arith in ("exact", "fast"):
        p = hh.Burgers1D(n, k, 1e-6, flux=flux, bc=1, xedges=xe, arith=arith)
        ode = hh.RkTvd(p, order=3)
        u = _t(u0, dev)
        t = ode.integrate(u, 0.0, 40 * dt, dt)
        assert t == ot and ode.fevals == fe
        ms = hh.MsTvd(p)
        um = _t(u0, dev)
        tm = ms.integrate(um, 0.0, 30 * dt, dt)
        assert tm == omt and ms.fevals == ofe
        if arith == "exact":
            assert np.array_equal(_h(u), ou)
            assert np.array_equal(_h(um), om)
        else:
            assert_close(_h(u), ou)
            assert_close(_h(um), om)


@pytest.mark.parametrize("arith", ["exact", "fast"])
def test_example1_godunov_zero_gradient(oracle, hh, dev, arith):
    """example/example1_burgers_1d_fv.f90: Godunov, fedges(0)=fedges(1), fedges(n)=fedges(n-1),
    limited linear initial profile on [-5, 5], dt = 1e-2 (100 cells)."""
    from hrweno_hip import workloads as W
    n = 100
    e, xc, w = W.linear_centres(-5.0, 5.0, n)
    u0 = np.clip(1.0 + (-0.5 - 1.0) / (2.0 + 4.0) * (xc + 4.0), -0.5, 1.0)
    u, t, ode = run_burgers(hh, dev, u0, n=n, flux=1, bc=1, tout=3.0, dt=1e-2, arith=arith,
                            xmin=-5.0, xmax=5.0)
    ou, ot, steps, fe = oracle.burgers_rk(u0, w, flux=1, bc=1, tout=3.0, dt=1e-2)
    assert t == ot and ode.fevals == fe
    if arith == "exact":
        assert np.array_equal(u, ou)
    else:
        assert_close(u, ou)


def test_itask_single_steps_and_state(oracle, hh, dev):
    from hrweno_hip import workloads as W
    c = W.config1(n=300, seed=5)
    p = hh.Burgers1D(c.n, 3, 1e-6, arith="exact")
    ode = hh.RkTvd(p, order=3)
    assert ode.istate == 1 and ode.fevals == 0 and ode.neq == 300 and ode.order == 3
    u = _t(c.u0, dev)
    t = 0.0
    for _ in range(5):
        t = ode.integrate(u, t, 1.0, c.dt, itask=2)
    assert ode.istate == 2 and ode.fevals == 15
    ou, ot, steps, fe = oracle.burgers_rk(c.u0, c.width, tout=4.5 * c.dt, dt=c.dt)
    assert steps == 5 and t == ot and np.array_equal(_h(u), ou)
    # already past tout: returns at once (src/hrweno_tvdode.f90:127)
    t2 = ode.integrate(u, t, t - 1.0, c.dt)
    assert t2 == t and ode.fevals == 15


def test_mstvd_burgers_exact(oracle, hh, dev):
    from hrweno_hip import workloads as W
    c = W.config1(n=500, seed=6)
    u, t, ode = run_burgers(hh, dev, c.u0, n=c.n, tout=0.1, dt=c.dt, method="ms")
    ou, ot, fe = oracle.burgers_ms(c.u0, c.width, tout=0.1, dt=c.dt)
    assert t == ot and ode.fevals == fe
    assert np.array_equal(u, ou)


@pytest.mark.parametrize("r", [4, 8])
def test_fast_mstvd_and_rhs(oracle, hh, dev, r):
    """FAST multistep (two pointwise operands per cell) and rhs-only passes of the 1D kernel.

    The reference's 5-step multistep method amplifies rounding on this WENO5-LF Burgers line
    at CFL 0.4: in the oracle itself a few-ulp change of u0 grows to 2e-13 after 20 steps and
    2e-9 after 100.  So the FAST form is compared over 4 starting + 10 multistep steps, where
    the problem is still well conditioned (the EXACT form is bitwise at any horizon)."""
    from hrweno_hip import workloads as W
    c = W.config1(n=3000, seed=12)
    p = hh.Burgers1D(c.n, 3, 1e-6, arith="fast").tune(0, r)
    ode = hh.MsTvd(p)
    u = _t(c.u0, dev)
    tout = W.steps_tout(0.0, c.dt, 14)
    t = ode.integrate(u, 0.0, tout, c.dt)
    ou, ot, fe = oracle.burgers_ms(c.u0, c.width, tout=tout, dt=c.dt)
    assert t == ot and ode.fevals == fe
    assert_close(_h(u), ou)
    # L = -(F+ - F-)/w cancels ~1000x at this resolution, so a flux-level rounding difference
    # is ~1e-12 of L itself on smooth cells: L is held to 1e-12 of max|L| (the state, above,
    # also to 1e-12 relative per cell)
    L = _h(p.rhs(0.0, _t(c.u0, dev)))
    oL = oracle.burgers_rhs(c.u0, c.width)
    assert np.max(np.abs(L - oL)) <= 1e-12 * np.abs(oL).max()


@pytest.mark.parametrize("k", [2, 3])
def test_persistent_form_rhs_and_mstvd_identical(hh, dev, k):
    """The persistent form (with 4 cells per lane: two chunks of loads in flight per wave)
    against one wavefront per chunk, bit for bit, in its three store patterns: an rhs-only pass
    (stores L), the multistep integrator (its starting RK3 stages store u; its multistep stages
    store u and L) on lengths around the first and the last interior chunk."""
    from hrweno_hip import workloads as W
    for n in (255, 256, 1000, 20000, 300001):
        c = W.config1(n=n, seed=30 + n)
        out = {}
        for variant in (0, 2):
            p = hh.Burgers1D(n, k, 1e-6, arith="fast").tune(variant, 4)
            L = _h(p.rhs(0.0, _t(c.u0, dev)))
            ode = hh.MsTvd(p)
            u = _t(c.u0, dev)
            ode.integrate(u, 0.0, W.steps_tout(0.0, c.dt, 8), c.dt)
            out[variant] = (L, _h(u))
        assert np.array_equal(out[0][0], out[2][0]), n
        assert np.array_equal(out[0][1], out[2][1]), n


def test_mstvd_resume_across_calls(oracle, hh, dev):
    """History persists across integrate calls (istate 2), as in the reference."""
    from hrweno_hip import workloads as W
    c = W.config1(n=400, seed=8)
    p = hh.Burgers1D(c.n, 3, 1e-6, arith="exact")
    ode = hh.MsTvd(p)
    u = _t(c.u0, dev)
    t = 0.0
    for tout in (0.02, 0.05, 0.08):
        t = ode.integrate(u, t, tout, c.dt)
    ou, ot, fe = oracle.burgers_ms(c.u0, c.width, tout=0.08, dt=c.dt)
    assert t == ot and ode.fevals == fe
    assert np.array_equal(_h(u), ou)


def test_problem_rhs(oracle, hh, dev):
    """One RHS evaluation: u0 + 1.0*L equals the oracle's RK1 step with dt = 1."""
    from hrweno_hip import workloads as W
    c = W.config1(n=777, seed=9)
    p = hh.Burgers1D(c.n, 3, 1e-6, arith="exact")
    L = _h(p.rhs(0.0, _t(c.u0, dev)))
    ou, *_ = oracle.burgers_rk(c.u0, c.width, order=1, tout=0.5, dt=1.0, itask=2)
    assert np.array_equal(c.u0 + 1.0 * L, ou)


def test_conservation_and_constant_state(hh, dev):
    n = 4096
    e = -1.0 + (2.0 / n) * np.arange(n + 1)
    w = e[1:] - e[:-1]
    xc = (e[:-1] + e[1:]) / 2
    u0 = 0.5 + np.sin(np.pi * xc) + (xc > 0.2)
    dt = 0.4 * (2.0 / n) / np.abs(u0).max()
    for arith in ("exact", "fast"):
        u, t, _ = run_burgers(hh, dev, u0, n=n, tout=0.2, dt=dt, arith=arith)
        assert abs(np.sum(u * w) - np.sum(u0 * w)) <= 1e-12 * np.sum(np.abs(u0) * w)
        c, *_ = run_burgers(hh, dev, np.full(n, 0.3), n=n, tout=0.01, dt=dt, arith=arith)
        assert np.all(c == 0.3)


def test_config2_full_size_exact_steps(oracle, hh, dev):
    """2**24 cells: 3 RK3 steps bit-identical to the oracle (OpenMP-threaded checker)."""
    from hrweno_hip import workloads as W
    c = W.config2()
    tout = W.steps_tout(0.0, c.dt, 3)
    u, t, ode = run_burgers(hh, dev, c.u0, n=c.n, tout=tout, dt=c.dt)
    oracle.set_threads(16)
    ou, ot, steps, fe = oracle.burgers_rk(c.u0, c.width, tout=tout, dt=c.dt)
    oracle.set_threads(1)
    assert steps == 3 and t == ot and ode.fevals == 9
    assert np.array_equal(u, ou)


def test_config2_full_size_fast_properties(hh, dev):
    """2**24 cells, 100 fast RK3 steps: mass conserved; agrees with EXACT to 1e-12."""
    from hrweno_hip import workloads as W
    c = W.config2()
    tout = W.steps_tout(0.0, c.dt, 100)
    uf, tf, of = run_burgers(hh, dev, c.u0, n=c.n, tout=tout, dt=c.dt, arith="fast")
    ue, te, oe = run_burgers(hh, dev, c.u0, n=c.n, tout=tout, dt=c.dt, arith="exact")
    assert of.fevals == 300 and tf == te
    m0 = np.sum(c.u0 * c.width)
    assert abs(np.sum(uf * c.width) - m0) <= 1e-12 * np.sum(np.abs(c.u0) * c.width)
    assert_close(uf, ue)


# ------------------------------------------------------------------------------ 2D PBE
def pbe_case(n1, n2, seed=4):
    from hrweno_hip import workloads as W
    rng = W.Rng(seed)
    cx, cy = rng.uniform(2, 0.2, 0.4)
    e1, c1, w1 = W.linear_centres(0.0, 1.0, n1)
    e2, c2, w2 = W.linear_centres(0.0, 1.0, n2)
    X1, X2 = np.meshgrid(c1, c2, indexing="xy")
    u0 = np.exp(-((X1 - cx) ** 2 + (X2 - cy) ** 2) / (2 * 0.05 ** 2)).ravel()
    dt = 0.4 / (1.5 * n1 + 0.5 * n2)
    return u0, e1, e2, w1, w2, dt


@pytest.mark.parametrize("bc", [2, 3])
@pytest.mark.parametrize("k", [1, 2, 3])
def test_pbe_mstvd_exact_bitwise(oracle, hh, dev, bc, k):
    n1, n2 = 131, 70        # n1 spans three 62-column strips, n2 several row blocks
    u0, e1, e2, w1, w2, dt = pbe_case(n1, n2)
    p = hh.PBE2D(n1, n2, k=k, bc=bc, arith="exact")
    ode = hh.MsTvd(p)
    u = _t(u0, dev)
    t = ode.integrate(u, 0.0, 15.5 * dt, dt)
    ou, ot, fe = oracle.pbe_ms(u0, e1, e2, k=k, bc=0 if bc == 2 else 1, tout=15.5 * dt, dt=dt)
    assert t == ot and ode.fevals == fe == 12 + 12
    assert np.array_equal(_h(u), ou)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_pbe_rktvd_exact_bitwise(oracle, hh, dev, order):
    n1, n2 = 64, 63
    u0, e1, e2, w1, w2, dt = pbe_case(n1, n2, seed=5)
    ode = hh.RkTvd(hh.PBE2D(n1, n2, bc=3, arith="exact"), order=order)
    u = _t(u0, dev)
    t = ode.integrate(u, 0.0, 6.5 * dt, dt)
    ou, ot, fe = oracle.pbe_ms(u0, e1, e2, bc=1, tout=6.5 * dt, dt=dt, order=order)
    assert t == ot and ode.fevals == fe
    assert np.array_equal(_h(u), ou)


def test_pbe_fast_within_tolerance_and_mass(oracle, hh, dev):
    n1, n2 = 200, 150
    u0, e1, e2, w1, w2, dt = pbe_case(n1, n2, seed=7)
    ode = hh.MsTvd(hh.PBE2D(n1, n2, bc=2, arith="fast"))
    u = _t(u0, dev)
    t = ode.integrate(u, 0.0, 40.5 * dt, dt)
    ou, ot, fe = oracle.pbe_ms(u0, e1, e2, bc=0, tout=40.5 * dt, dt=dt)
    g = _h(u)
    assert_close(g, ou)
    mass = lambda q: np.sum(q.reshape(n2, n1) * w1[None, :] * w2[:, None])  # noqa: E731
    assert abs(mass(g) - mass(u0)) <= 1e-12 * mass(u0)


@pytest.mark.parametrize("n1", [1, 2, 3, 59, 60, 61, 119, 120, 121])
def test_pbe_fast_k3_strip_edges(oracle, hh, dev, n1):
    """FAST, k = 3: strips of 60 output columns whose lanes take their neighbours' smoothness
    triples by lane shifts; grids that end just before, on and after a strip boundary stay
    within the FAST tolerance of the oracle."""
    n2 = 9
    u0, e1, e2, w1, w2, dt = pbe_case(n1, n2, seed=13)
    ode = hh.MsTvd(hh.PBE2D(n1, n2, k=3, bc=3, arith="fast"))
    u = _t(u0, dev)
    t = ode.integrate(u, 0.0, 9.5 * dt, dt)
    ou, ot, fe = oracle.pbe_ms(u0, e1, e2, k=3, bc=1, tout=9.5 * dt, dt=dt)
    assert t == ot and ode.fevals == fe
    assert_close(_h(u), ou)


@pytest.mark.parametrize("n1,n2", [(1, 1), (2, 3), (5, 4), (63, 2), (125, 9)])
@pytest.mark.parametrize("k", [1, 2, 3])
def test_pbe_small_and_ragged_exact_bitwise(oracle, hh, dev, n1, n2, k):
    """Grids narrower than the x1 window, one row, a strip plus one column: the lane-shifted
    x1 windows clamp at the edges exactly as the reference's ghost-free stencils do."""
    u0, e1, e2, w1, w2, dt = pbe_case(n1, n2, seed=11)
    ode = hh.MsTvd(hh.PBE2D(n1, n2, k=k, bc=3, arith="exact"))
    u = _t(u0, dev)
    t = ode.integrate(u, 0.0, 7.5 * dt, dt)
    ou, ot, fe = oracle.pbe_ms(u0, e1, e2, k=k, bc=1, tout=7.5 * dt, dt=dt)
    assert t == ot and ode.fevals == fe
    assert np.array_equal(_h(u), ou)


@pytest.mark.parametrize("arith", ["exact", "fast"])
def test_pbe_rows_per_wave_identical(hh, dev, arith):
    """The rows each wavefront marches over (tuned or automatic) do not change any result."""
    n1, n2 = 131, 70
    u0, e1, e2, w1, w2, dt = pbe_case(n1, n2, seed=9)
    res = []
    for ty in (0, 1, 5, 16, 64, 100):
        ode = hh.MsTvd(hh.PBE2D(n1, n2, bc=3, arith=arith).tune(0, ty))
        u = _t(u0, dev)
        ode.integrate(u, 0.0, 10.5 * dt, dt)
        res.append(_h(u))
    for r in res[1:]:
        assert np.array_equal(r, res[0])


def test_config4_full_size_exact_steps(oracle, hh, dev):
    """2048 x 2048: 4 starting RK3 steps + 2 multistep steps bit-identical to the oracle."""
    from hrweno_hip import workloads as W
    c = W.config4()
    e, _, _ = W.linear_centres(0.0, 1.0, c.n1)
    tout = W.steps_tout(0.0, c.dt, 6)
    ode = hh.MsTvd(hh.PBE2D(c.n1, c.n2, bc=3, arith="exact"))
    u = _t(c.u0, dev)
    t = ode.integrate(u, 0.0, tout, c.dt)
    oracle.set_threads(16)
    ou, ot, fe = oracle.pbe_ms(c.u0, e, e, bc=1, tout=tout, dt=c.dt)
    oracle.set_threads(1)
    assert t == ot and ode.fevals == fe == 14
    assert np.array_equal(_h(u), ou)


def test_config4_full_size_exact_500_steps_and_fast(oracle, hh, dev):
    """The benchmarked 2D workload at its full size and horizon: 2048 x 2048, the multistep
    method for 500 steps (4 RK3 starting steps + 496), zero inflow / upwind outflow.  The EXACT
    form (the one timed for this workload) is bit-identical to the oracle over all 500 steps.
    The FAST form stays within 1e-12 of it (per cell of max|u|, and relative on every cell
    above 1e-3 max|u|) over the first 20 steps; beyond that the reference's own method
    amplifies rounding (a 1-ulp change of u0 is ~1e-11 by step 200 and ~1e-1 by step 500 in
    the reference's own arithmetic: test_oracle_tvd.py::test_pbe_multistep_amplifies_rounding),
    so no other rounding can stay within 1e-12 of the reference over the whole horizon."""
    from hrweno_hip import workloads as W
    c = W.config4()
    e, _, _ = W.linear_centres(0.0, 1.0, c.n1)
    res = {}
    for steps in (20, c.steps):
        tout = W.steps_tout(0.0, c.dt, steps)
        for arith in ("exact", "fast"):
            ode = hh.MsTvd(hh.PBE2D(c.n1, c.n2, bc=3, arith=arith))
            u = _t(c.u0, dev)
            t = ode.integrate(u, 0.0, tout, c.dt)
            assert ode.fevals == 12 + steps - 4
            res[(arith, steps)] = (_h(u), t)
    assert res[("fast", 20)][1] == res[("exact", 20)][1]
    assert_close(res[("fast", 20)][0], res[("exact", 20)][0])
    oracle.set_threads(16)
    try:
        ou, ot, fe = oracle.pbe_ms(c.u0, e, e, bc=1, tout=W.steps_tout(0.0, c.dt, c.steps), dt=c.dt)
    finally:
        oracle.set_threads(1)
    assert res[("exact", c.steps)][1] == ot and fe == 12 + c.steps - 4
    assert np.array_equal(res[("exact", c.steps)][0], ou)


# ------------------------------------------------------------------------------ callbacks
def test_callback_rk_linear_ode_closed_form(hh, dev):
    """rktvd(fu, neq, order) with a device callback fu = a*u (test/test_tvdode.f90)."""
    import torch
    a = _t(-1.0 + np.arange(10) * 4.0 / 9, dev)
    A = a.cpu().numpy()

    for order in (1, 2, 3):
        def fu(t, up, udp, stream):
            from hrweno_hip import _lib
            # udot = LF((0 + 1*a) u, ...) with vm = vp = u equals a*u exactly
            return _lib.lib().hrw_flux_lax_friedrichs(1, 0.0, 1.0, 10, up, up, a.data_ptr(), t,
                                                      0.0, udp, stream)
        ode = hh.RkTvd(fu, neq=10, order=order)
        u = _t(np.full(10, 0.1), dev)
        t0, tout = -1.3, 1.5
        dt = (tout - t0) / 3000
        t = ode.integrate(u, t0, tout, dt)
        torch.cuda.synchronize()
        nsteps = ode.fevals // order
        z = A * dt
        R = {1: 1 + z, 2: 1 + z + z ** 2 / 2, 3: 1 + z + z ** 2 / 2 + z ** 3 / 6}[order]
        got = _h(u)
        assert np.allclose(got, 0.1 * R ** nsteps, rtol=1e-10, atol=0)
        rtol = [2e-2, 1e-3, 1e-3][order - 1]
        assert np.allclose(got, 0.1 * np.exp(A * (t - t0)), rtol=rtol, atol=0)


def test_callback_ms_linear_ode(oracle, hh, dev):
    a_np = -1.0 + np.arange(10) * 4.0 / 9
    a = _t(a_np, dev)

    def fu(t, up, udp, stream):
        from hrweno_hip import _lib
        return _lib.lib().hrw_flux_lax_friedrichs(1, 0.0, 1.0, 10, up, up, a.data_ptr(), t, 0.0,
                                                  udp, stream)
    ode = hh.MsTvd(fu, neq=10, arith="exact")
    u = _t(np.full(10, 0.1), dev)
    t = ode.integrate(u, -1.3, 1.5, 1e-3)
    ou, ot, fe = oracle.linear_ode(a_np, np.full(10, 0.1), 0, -1.3, 1.5, 1e-3)
    assert t == ot and ode.fevals == fe
    assert np.array_equal(_h(u), ou)


def test_callback_burgers_equals_fused(hh, dev):
    """The generic callback path (fu = device RHS) gives the fused path's numbers."""
    from hrweno_hip import workloads as W
    c = W.config1(n=512, seed=13)
    p = hh.Burgers1D(c.n, 3, 1e-6, arith="exact")

    def fu(t, up, udp, stream):
        from hrweno_hip import _lib
        return _lib.lib().hrw_problem_rhs(p._h, t, up, udp, stream)
    ode_cb = hh.RkTvd(fu, neq=c.n, order=3, arith="exact")
    u1 = _t(c.u0, dev)
    t1 = ode_cb.integrate(u1, 0.0, 0.05, c.dt)
    u2, t2, ode = run_burgers(hh, dev, c.u0, n=c.n, tout=0.05, dt=c.dt)
    assert t1 == t2 and ode_cb.fevals == ode.fevals
    assert np.array_equal(_h(u1), u2)
