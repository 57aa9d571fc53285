"""Device discrete adjoint (hrw_adjoint_* through the C-ABI) against the oracle adjoint,
finite differences of the device forward, torch.autograd.gradcheck, the conservation closed
form, checkpoint-interval invariance and batched == per-line."""
import numpy as np
import pytest

from hrweno_hip import workloads as W

pytestmark = pytest.mark.gpu
# The reverse pass runs in the FAST arithmetic form (FMA, reciprocal + Newton) against the
# oracle's reference-order arithmetic: agreement to 1e-10 of the largest component.
TOL = 1e-10


def _case(n, seed, noise=0.0):
    e, xc, w = W.linear_centres(-1.0, 1.0, n)
    u0 = W.smooth_step_ic(xc, W.Rng(seed))
    if noise:
        u0 = u0 + noise * np.random.default_rng(seed).standard_normal(n)
    dt = W.cfl_dt(u0, 2.0 / n)
    return u0, w, dt


def _dev(a, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev).clone()


@pytest.mark.parametrize("n", [1, 2, 3, 5, 64, 200, 497, 1000, 4097])
@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("order", [1, 3])
def test_adjoint_matches_oracle(oracle, hh, dev, n, k, order):
    u0, w, dt = _case(n, 100 + n + k)
    steps = 3
    gN = np.random.default_rng(n).standard_normal(n)
    ou, og = oracle.burgers_rk_adjoint(u0, gN, w, nsteps=steps, dt=dt, k=k, order=order)
    adj = hh.Adjoint(hh.Burgers1D(n, k=k), nsteps=steps, dt=dt, order=order, ckpt_every=2)
    uN = adj.forward(_dev(u0, dev))
    g0 = adj.backward(_dev(gN, dev)).cpu().numpy()
    assert np.max(np.abs(uN.cpu().numpy() - ou)) <= 1e-12 * max(np.abs(ou).max(), 1.0)
    assert np.max(np.abs(g0 - og)) <= TOL * np.abs(og).max(), (n, k, order)


@pytest.mark.parametrize("alpha", [-1.0, 2.5])
def test_adjoint_order2_and_fixed_alpha(oracle, hh, dev, alpha):
    n, steps = 777, 6
    u0, w, dt = _case(n, 9)
    gN = np.cos(np.arange(n))
    ou, og = oracle.burgers_rk_adjoint(u0, gN, w, nsteps=steps, dt=dt, order=2, alpha=alpha)
    adj = hh.Adjoint(hh.Burgers1D(n, alpha=alpha), nsteps=steps, dt=dt, order=2)
    adj.forward(_dev(u0, dev))
    g0 = adj.backward(_dev(gN, dev)).cpu().numpy()
    assert np.max(np.abs(g0 - og)) <= TOL * np.abs(og).max()


def test_checkpoint_interval_is_invisible(hh, dev):
    """Recomputed segments reproduce the recorded ones bit for bit."""
    n, steps = 3000, 23
    u0, w, dt = _case(n, 4)
    gN = _dev(np.sin(np.arange(n) * 0.1), dev)
    outs = []
    for ck in (1, 5, 23, 100, 0):
        adj = hh.Adjoint(hh.Burgers1D(n), nsteps=steps, dt=dt, ckpt_every=ck)
        if ck == 0:   # automatic: the whole run fits, every stage input is kept
            assert adj.info()["ckpt_every"] == steps
        adj.forward(_dev(u0, dev))
        before = adj.info()["device_bytes"]
        outs.append(adj.backward(gN).cpu().numpy())
        # with more than one segment the backward takes a second set of stage inputs and
        # recomputes each segment on a second stream beside the reverse of the next
        grown = adj.info()["device_bytes"] > before
        assert grown == (ck < steps and ck != 0), ck
        # a second backward of the same forward (segments recomputed again) is identical
        outs.append(adj.backward(gN).cpu().numpy())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


@pytest.mark.parametrize("k", [1, 2, 3])
def test_cells_per_lane_identical(hh, dev, k):
    """The reverse kernel with 2 or 4 cells per lane agrees to rounding: a cell's window
    cotangents from the same lane and from the neighbouring lanes are added in an order that
    depends on the lane mapping, so the last bits may differ."""
    for n in (3, 250, 4099, 70001):
        u0, w, dt = _case(n, 40 + n)
        gN = _dev(np.sin(np.arange(n) * 0.7), dev)
        outs = []
        for r in (4, 2):
            adj = hh.Adjoint(hh.Burgers1D(n, k=k), nsteps=5, dt=dt, ckpt_every=2).set_cells_per_lane(r)
            adj.forward(_dev(u0, dev))
            outs.append(adj.backward(gN).cpu().numpy())
        assert np.max(np.abs(outs[0] - outs[1])) <= 1e-13 * np.abs(outs[0]).max(), n


def test_conserved_total_full_size(hh, dev):
    """Periodic FV conserves sum_i w_i u_i, so the gradient of that total is w exactly:
    size-independent closed form at 2^20 cells."""
    import torch
    n, steps = 1 << 20, 40
    u0, w, dt = _case(n, 5)
    adj = hh.Adjoint(hh.Burgers1D(n), nsteps=steps, dt=dt, ckpt_every=20)
    adj.forward(_dev(u0, dev))
    g0 = adj.backward(_dev(w, dev))
    torch.cuda.synchronize()
    assert np.max(np.abs(g0.cpu().numpy() - w)) <= 1e-12 * np.abs(w).max()


def test_fd_64_cells_full_size(hh, dev):
    """The adjoint-form workload: 2^20 cells, 400 RK3 steps, checkpoints every 20; loss
    0.5*|u_N - target|^2 with the target run from u0 + N(0, 0.01); the gradient at 64 random
    cells against Richardson-extrapolated central differences of the device forward,
    (4 D(h) - D(2h))/3 with h = 2e-6, whose truncation is O(h^4) (a plain central difference
    keeps an O(h^2) term that reaches 1e-5 relative at h = 1e-6 where the eps = 1e-6 WENO
    weights turn sharply).  The rounding noise of the differences falls as 1/h: with
    h = 1e-6 it reached 1.3e-6 relative at a cell whose gradient is 1.4 % of the largest, and
    8e-9 with h = 2e-6, where the two pairs agree to ~1e-8 at the smooth cells (the O(h^4)
    term is below that).  Every cell agrees to 1e-6 relative, except by rule a cell whose
    +-h perturbation moves the argmax of |u| (of the initial or of the final state): there
    alpha = max|u| has a kink, the loss is not differentiable, and the adjoint returns the
    subgradient of the first cell attaining the maximum."""
    import torch
    c, noise = W.config5()
    n, steps, h, u0, dt = c.n, c.steps, 2e-6, c.u0, c.dt
    adj = hh.Adjoint(hh.Burgers1D(n), nsteps=steps, dt=dt, ckpt_every=20)
    target = adj.forward(_dev(u0 + noise, dev)).clone()
    u0d = _dev(u0, dev)
    uN = adj.forward(u0d).clone()
    g0 = adj.backward(uN - target).cpu().numpy()
    cells = np.random.default_rng(5).choice(n, 64, replace=False)
    am0, amN = int(torch.argmax(torch.abs(u0d))), int(torch.argmax(torch.abs(uN)))

    def fd(j, step):
        """Central difference of the loss, and whether either run moved an argmax of |u|."""
        runs = []
        moved = False
        for sgn in (1.0, -1.0):
            up = u0d.clone()
            up[j] += sgn * step
            moved |= int(torch.argmax(torch.abs(up))) != am0
            a = adj.forward(up)
            moved |= int(torch.argmax(torch.abs(a))) != amN
            runs.append(a - target)
        ap, am = runs
        return float(torch.sum((ap - am) * (ap + am)).item()) / (4 * step), moved

    g = g0[cells]
    d1 = [fd(j, h) for j in cells]
    d2 = [fd(j, 2 * h) for j in cells]
    rich = np.array([(4 * a[0] - b[0]) / 3 for a, b in zip(d1, d2)])
    kink = np.array([a[1] or b[1] for a, b in zip(d1, d2)])
    plain = np.array([a[0] for a in d1])
    rel = np.abs(g - rich) / np.abs(rich)
    rel_plain = np.abs(g - plain) / np.abs(plain)
    print(f"64 cells, {int(kink.sum())} at an argmax kink; Richardson: max {rel[~kink].max():.2e} "
          f"median {np.median(rel[~kink]):.2e}; plain central h=2e-6: max "
          f"{rel_plain[~kink].max():.2e}")
    assert kink.sum() <= 2
    assert rel[~kink].max() <= 1e-6


def test_gradcheck_autograd(hh, dev):
    import torch
    from hrweno_hip.autograd import evolve
    for k, order in ((3, 3), (2, 1)):
        n = 64
        u0, w, dt = _case(n, 21 + k, noise=0.05)
        adj = hh.Adjoint(hh.Burgers1D(n, k=k), nsteps=8, dt=dt, order=order, ckpt_every=3)
        x = _dev(u0, dev).requires_grad_(True)
        assert torch.autograd.gradcheck(lambda v: evolve(v, adj), (x,), eps=1e-6, atol=1e-7,
                                        rtol=1e-5, nondet_tol=0.0)


def test_autograd_loss_matches_backward(hh, dev):
    import torch
    from hrweno_hip.autograd import evolve
    n = 5000
    u0, w, dt = _case(n, 8)
    adj = hh.Adjoint(hh.Burgers1D(n), nsteps=30, dt=dt, ckpt_every=7)
    target = _dev(np.zeros(n), dev)
    x = _dev(u0, dev).requires_grad_(True)
    uN = evolve(x, adj)
    loss = 0.5 * ((uN - target) ** 2).sum()
    loss.backward()
    g_ref = adj.backward((uN.detach() - target)).cpu().numpy()
    assert np.array_equal(x.grad.cpu().numpy(), g_ref)


def test_batched_adjoint_equals_per_line(oracle, hh, dev):
    """256-line style ensemble (here 24 lines, k in {2,3}): the batched adjoint equals the
    single-line adjoint of each line bit for bit, and the oracle within TOL."""
    import torch
    rng = np.random.default_rng(3)
    lens = np.concatenate([[1, 5, 64, 1024], np.rint(np.exp(rng.uniform(np.log(1024), np.log(16384), 20)))]).astype(np.int64)
    ks = rng.integers(2, 4, lens.size)
    u0s, dts, widths = [], [], []
    for l, n in enumerate(lens):
        u0, w, dt = _case(int(n), 300 + l)
        u0s.append(u0)
        dts.append(dt)
        widths.append(w)
    steps = 7
    b = hh.BurgersBatch(lens, ks, np.array(dts))
    adj = hh.Adjoint(b, nsteps=steps, ckpt_every=3)
    u = b.pack(u0s, dev)
    uN = adj.forward(u)
    gNs = [np.cos(np.arange(n) * 0.3 + l) for l, n in enumerate(lens)]
    g0 = b.unpack(adj.backward(b.pack(gNs, dev)))
    outN = b.unpack(uN)
    for l, n in enumerate(lens):
        a1 = hh.Adjoint(hh.Burgers1D(int(n), k=int(ks[l])), nsteps=steps, dt=dts[l], ckpt_every=3)
        u1 = a1.forward(_dev(u0s[l], dev))
        g1 = a1.backward(_dev(gNs[l], dev))
        torch.cuda.synchronize()
        assert np.array_equal(outN[l], u1.cpu().numpy()), l
        assert np.array_equal(g0[l], g1.cpu().numpy()), l
        ou, og = oracle.burgers_rk_adjoint(u0s[l], gNs[l], widths[l], nsteps=steps, dt=dts[l],
                                           k=int(ks[l]))
        assert np.max(np.abs(g0[l] - og)) <= TOL * np.abs(og).max(), l


def test_config5_batch_full_equals_per_line(hh, dev):
    """The batched adjoint workload at its full size and horizon: 256 lines of 1024..16384
    cells, 400 RK3 steps, checkpoints every 20, each line's loss against its own perturbed
    target.  The batched forward and gradient equal each line's own single-line adjoint bit for
    bit."""
    import torch
    cases, noises = W.config5_batch()
    steps = cases[0].steps
    lens = np.array([c.n for c in cases])
    b = hh.BurgersBatch(lens, np.full(lens.size, 3), np.array([c.dt for c in cases]))
    adj = hh.Adjoint(b, nsteps=steps, ckpt_every=20)
    target = adj.forward(b.pack([c.u0 + z for c, z in zip(cases, noises)], dev)).clone()
    uN = adj.forward(b.pack([c.u0 for c in cases], dev)).clone()
    g0 = b.unpack(adj.backward(uN - target))
    outN, tN = b.unpack(uN), b.unpack(target)
    for l, c in enumerate(cases):
        a1 = hh.Adjoint(hh.Burgers1D(c.n), nsteps=steps, dt=c.dt, ckpt_every=20)
        u1 = a1.forward(_dev(c.u0, dev)).clone()
        g1 = a1.backward(u1 - _dev(tN[l], dev))
        torch.cuda.synchronize()
        assert np.array_equal(outN[l], u1.cpu().numpy()), l
        assert np.array_equal(g0[l], g1.cpu().numpy()), l


def test_adjoint_argument_errors(hh):
    # (zero-gradient ends and the Godunov flux take the general reverse pass; a non-uniform
    # grid's own reconstruction coefficients are not differentiated)
    e = np.linspace(-1.0, 1.0, 101) ** 3
    with pytest.raises(hh.HrwError, match="linear grid"):
        hh.Adjoint(hh.Burgers1D(100, bc=hh.BC_ZERO_GRADIENT, xedges=e), nsteps=3, dt=1e-3)
    with pytest.raises(hh.HrwError, match="ckpt_every"):
        hh.Adjoint(hh.Burgers1D(100), nsteps=3, dt=1e-3, ckpt_every=-1)
    adj = hh.Adjoint(hh.Burgers1D(100), nsteps=3, dt=1e-3)
    import torch
    with pytest.raises(hh.HrwError, match="recorded"):
        adj.backward(torch.zeros(100, dtype=torch.float64, device="cuda"))
