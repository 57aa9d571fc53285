"""Batched ensemble of uneven periodic Burgers lines (one launch per stage for all lines, with and
without a captured hipGraph) vs the one-launch-per-line path and vs the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-12


def ensemble(nlines, seed, nmax=3000, steps=6):
    from hrweno_hip import workloads as W
    rng = np.random.default_rng(seed)
    lens = np.concatenate([[1, 2, 5, 63, 64, 513], np.rint(np.exp(rng.uniform(0, np.log(nmax), nlines - 6)))])
    lens = lens.astype(np.int64)[:nlines]
    ks = rng.integers(1, 4, nlines)
    u0s, dts, widths = [], [], []
    for l, n in enumerate(lens):
        e, xc, w = W.linear_centres(-1.0, 1.0, int(n))
        u0 = W.smooth_step_ic(xc, W.Rng(seed * 1000 + l))
        u0s.append(u0)
        dts.append(W.cfl_dt(u0, 2.0 / n))
        widths.append(w)
    return lens, ks, np.array(dts), u0s, widths


def test_batched_graph_and_per_line_agree_and_match_oracle(oracle, hh, dev):
    from hrweno_hip import workloads as W
    lens, ks, dts, u0s, widths = ensemble(40, 3)
    steps = 6
    outs = {}
    for mode in ("batched", "graph", "per_line"):
        b = hh.BurgersBatch(lens, ks, dts)
        u = b.pack(u0s, dev)
        if mode == "per_line":
            b.step_per_line(u, steps)
        else:
            b.step(u, steps, graph=(mode == "graph"))
        import torch
        torch.cuda.synchronize()
        outs[mode] = b.unpack(u)
        # per-line time accumulates in FP64 like the reference loop
        for l in range(len(lens)):
            t = 0.0
            for _ in range(steps):
                t = t + dts[l]
            assert b.t[l] == t
    for l in range(len(lens)):
        assert np.array_equal(outs["batched"][l], outs["graph"][l])
        assert np.array_equal(outs["batched"][l], outs["per_line"][l])
        ou, ot, st, fe = oracle.burgers_rk(u0s[l], widths[l], k=int(ks[l]),
                                           tout=W.steps_tout(0.0, dts[l], steps), dt=dts[l])
        assert st == steps
        g = outs["batched"][l]
        scale = np.abs(ou).max()
        assert np.max(np.abs(g - ou)) <= TOL * scale, (l, lens[l], ks[l])


def test_tune_keeps_results(hh, dev):
    import torch
    lens, ks, dts, u0s, widths = ensemble(24, 5)
    res = []
    for form in (0, 1):
        for r in (4, 6, 8):
            b = hh.BurgersBatch(lens, ks, dts)
            u = b.pack(u0s, dev)
            b.set_cells_per_lane(r)
            b.set_form(form)
            assert b.config() == (r, form)
            b.step(u, 4, graph=True)
            torch.cuda.synchronize()
            res.append(u.cpu().numpy())
    for x in res[1:]:
        assert np.array_equal(res[0], x)
    b = hh.BurgersBatch(lens, ks, dts)
    u = b.pack(u0s, dev)
    assert b.tune(u) in (4, 6, 8)
    assert b.config()[1] in (0, 1)


@pytest.mark.parametrize("flux,alpha", [(0, -1.0), (0, 1.5), (1, -1.0)])
@pytest.mark.parametrize("form", [0, 1])
def test_forms_match_per_line_every_flux(hh, dev, flux, alpha, form):
    """Both batch kernels (one wavefront per chunk; persistent wavefronts with the next chunk's
    loads issued ahead, interior chunks first) give the per-line path's bits for every flux,
    on lines from 1 cell to a few dozen chunks."""
    import torch
    lens, ks, dts, u0s, widths = ensemble(30, 11, nmax=12000)
    for r in (4, 6, 8):
        outs = []
        for per_line in (False, True):
            b = hh.BurgersBatch(lens, ks, dts, flux=flux, alpha=alpha)
            b.set_cells_per_lane(r)
            b.set_form(form)
            u = b.pack(u0s, dev)
            if per_line:
                b.step_per_line(u, 3)
            else:
                b.step(u, 3, graph=True)
            torch.cuda.synchronize()
            outs.append(u)
        assert torch.equal(outs[0], outs[1]), (r, form)


def test_config3_subset(oracle, hh, dev):
    """64 lines of the 4096-line ensemble (lengths up to 65536): batched == per-line bitwise,
    and four lines against the oracle."""
    import torch
    from hrweno_hip import workloads as W
    cases = W.config3(nlines=64, steps=10)
    lens = np.array([c.n for c in cases])
    ks = np.array([c.k for c in cases])
    dts = np.array([c.dt for c in cases])
    b1 = hh.BurgersBatch(lens, ks, dts)
    u1 = b1.pack([c.u0 for c in cases], dev)
    b1.step(u1, 10, graph=True)
    b2 = hh.BurgersBatch(lens, ks, dts)
    u2 = b2.pack([c.u0 for c in cases], dev)
    b2.step_per_line(u2, 10)
    b3 = hh.BurgersBatch(lens, ks, dts)
    b3.set_form(1)
    u3 = b3.pack([c.u0 for c in cases], dev)
    b3.step(u3, 10, graph=True)
    torch.cuda.synchronize()
    assert torch.equal(u1, u2)
    assert torch.equal(u3, u2)
    out = b1.unpack(u1)
    for l in (0, 7, 31, 63):
        c = cases[l]
        ou, ot, st, fe = oracle.burgers_rk(c.u0, c.width, k=c.k, tout=W.steps_tout(0.0, c.dt, 10),
                                           dt=c.dt)
        scale = np.abs(ou).max()
        assert np.max(np.abs(out[l] - ou)) <= TOL * scale


def test_config3_full_ensemble(oracle, hh, dev):
    """The benchmarked ensemble at its full size and horizon: 4096 lines of 256..65536 cells,
    k in {2, 3}, 200 RK3 steps.  The batched path (autotuned form and width, graph replay)
    equals the one-launch-per-line path bit for bit on every line, and 16 lines spread over the
    length range stay within the FAST tolerance of the oracle."""
    import torch
    from hrweno_hip import workloads as W
    steps = 200
    cases = W.config3(steps=steps)
    lens = np.array([c.n for c in cases])
    ks = np.array([c.k for c in cases])
    dts = np.array([c.dt for c in cases])
    b = hh.BurgersBatch(lens, ks, dts)
    u = b.pack([c.u0 for c in cases], dev)
    b.tune(u)
    b.step(u, steps, graph=True)
    bp = hh.BurgersBatch(lens, ks, dts)
    up = bp.pack([c.u0 for c in cases], dev)
    bp.step_per_line(up, steps)
    torch.cuda.synchronize()
    assert torch.equal(u, up)
    assert np.array_equal(b.t, bp.t)
    out = b.unpack(u)
    order = np.argsort(lens)
    oracle.set_threads(16)
    try:
        for l in order[np.linspace(0, lens.size - 1, 16).astype(int)]:
            c = cases[l]
            ou, ot, st, fe = oracle.burgers_rk(c.u0, c.width, k=c.k,
                                               tout=W.steps_tout(0.0, c.dt, steps), dt=c.dt)
            assert st == steps and ot == b.t[l]
            scale = np.abs(ou).max()
            d = np.abs(out[l] - ou)
            assert d.max() <= TOL * scale, (l, c.n, d.max() / scale)
            big = np.abs(ou) >= 1e-3 * scale
            assert np.max(d[big] / np.abs(ou[big])) <= TOL, (l, c.n)
    finally:
        oracle.set_threads(1)


@pytest.mark.parametrize("form", [0, 1])
def test_few_long_lines_match_per_line(hh, dev, form):
    """A few long lines: whole blocks of wavefronts on one line, whose line maxima are reduced in
    the block before the line's slot is updated; every alpha and so every bit as per line."""
    import torch
    from hrweno_hip import workloads as W
    lens = np.array([65536, 70001, 131071])
    ks = np.array([3, 2, 3])
    u0s, dts = [], []
    for l, n in enumerate(lens):
        e, xc, w = W.linear_centres(-1.0, 1.0, int(n))
        u0 = W.smooth_step_ic(xc, W.Rng(500 + l))
        u0s.append(u0)
        dts.append(W.cfl_dt(u0, 2.0 / n))
    dts = np.array(dts)
    outs = []
    for per_line in (False, True):
        b = hh.BurgersBatch(lens, ks, dts)
        b.set_form(form)
        u = b.pack(u0s, dev)
        if per_line:
            b.step_per_line(u, 4)
        else:
            b.step(u, 4, graph=True)
        torch.cuda.synchronize()
        outs.append(u)
    assert torch.equal(outs[0], outs[1])
