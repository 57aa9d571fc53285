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
        outs.append(adj.backward(gN).cpu().numpy())
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
    0.5*|u_N - target|^2 with the target run from u0 + N(0, 0.01); central differences
    (h = 1e-6) of the device forward at 64 random cells.

    The 64-cell gradient agrees with the differences to 1e-6 in norm and in median per cell.
    Single cells can differ more: the central difference's O(h^2) truncation is large where
    the eps = 1e-6 WENO weights turn sharply (they vary on a u-scale of ~3e-4; measured
    1e-5 at the worst of 64 cells, 2.6e-3 at h = 1e-5), so the worst cell is checked to move
    toward the adjoint when h is halved."""
    import torch
    c, noise = W.config5()
    n, steps, h, u0, dt = c.n, c.steps, 1e-6, c.u0, c.dt
    adj = hh.Adjoint(hh.Burgers1D(n), nsteps=steps, dt=dt, ckpt_every=20)
    target = adj.forward(_dev(u0 + noise, dev)).clone()
    u0d = _dev(u0, dev)
    uN = adj.forward(u0d)
    g0 = adj.backward(uN - target).cpu().numpy()
    cells = np.random.default_rng(5).choice(n, 64, replace=False)

    def fd(j, step):
        up = u0d.clone()
        up[j] += step
        um = u0d.clone()
        um[j] -= step
        ap = adj.forward(up) - target
        am = adj.forward(um) - target
        return float(torch.sum((ap - am) * (ap + am)).item()) / (4 * step)

    f1 = np.array([fd(j, h) for j in cells])
    g = g0[cells]
    rel = np.abs(g - f1) / np.abs(f1)
    norm = np.linalg.norm(g - f1) / np.linalg.norm(f1)
    w = int(np.argmax(rel))
    half = abs(g[w] - fd(cells[w], h / 2)) / abs(f1[w])
    print(f"central h=1e-6 on 64 cells: norm {norm:.2e} median {np.median(rel):.2e} "
          f"max {rel.max():.2e} (cell {cells[w]}; {half:.2e} at h/2)")
    assert norm <= 1e-6 and np.median(rel) <= 1e-6
    assert rel.max() <= 5e-5 and half < rel[w]


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


def test_adjoint_argument_errors(hh):
    with pytest.raises(hh.HrwError, match="periodic"):
        hh.Adjoint(hh.Burgers1D(100, bc=hh.BC_ZERO_GRADIENT), nsteps=3, dt=1e-3)
    with pytest.raises(hh.HrwError, match="ckpt_every"):
        hh.Adjoint(hh.Burgers1D(100), nsteps=3, dt=1e-3, ckpt_every=-1)
    adj = hh.Adjoint(hh.Burgers1D(100), nsteps=3, dt=1e-3)
    import torch
    with pytest.raises(hh.HrwError, match="recorded"):
        adj.backward(torch.zeros(100, dtype=torch.float64, device="cuda"))
