"""Device reconstruction and fluxes vs the oracle (through the C-ABI).

Tolerances: EXACT arithmetic must equal the oracle bit for bit (numeric equality);
FAST arithmetic must agree to 1e-12 relative to max|v| per cell (FMA reordering and the
common-denominator weights only).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _t(a, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)


def _h(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def lines(rng):
    for n in (1, 2, 3, 5, 64, 255, 256, 257, 1000, 4097):
        v = rng.standard_normal(n) * 10.0 ** rng.uniform(-6, 3)
        if n % 2:
            v = np.round(v * 4) / 4       # plateaus and jumps
        yield v


@pytest.mark.parametrize("k", [1, 2, 3])
def test_reconstruct_exact_bitwise(oracle, hh, dev, k):
    rng = np.random.default_rng(100 + k)
    for v in lines(rng):
        w = hh.Weno(v.size, k, 1e-6, arith="exact")
        vl, vr = w.reconstruct(_t(v, dev))
        ol, orr = oracle.reconstruct(v, k, 1e-6)
        assert np.array_equal(_h(vl), ol) and np.array_equal(_h(vr), orr), (k, v.size)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_reconstruct_fast_within_tolerance(oracle, hh, dev, k):
    rng = np.random.default_rng(200 + k)
    for v in lines(rng):
        w = hh.Weno(v.size, k, 1e-6, arith="fast")
        vl, vr = w.reconstruct(_t(v, dev))
        ol, orr = oracle.reconstruct(v, k, 1e-6)
        scale = np.abs(v).max()
        assert np.max(np.abs(_h(vl) - ol)) <= TOL * scale
        assert np.max(np.abs(_h(vr) - orr)) <= TOL * scale


@pytest.mark.parametrize("arith", ["exact", "fast"])
@pytest.mark.parametrize("k", [1, 2, 3])
def test_reference_pulse(hh, dev, k, arith):
    """test/test_hrweno.f90:29-67 on the device."""
    v = np.zeros(30)
    v[9:20] = 1.0
    vl, vr = hh.Weno(30, k, 1e-6, arith=arith).reconstruct(_t(v, dev))
    assert np.max(np.abs(_h(vl) - v)) < 1e-6 and np.max(np.abs(_h(vr) - v)) < 1e-6


def test_reconstruct_strided_section(oracle, hh, dev):
    """Strided sections reach reconstruct (example2:107)."""
    import torch
    rng = np.random.default_rng(7)
    a = rng.standard_normal((50, 7))
    A = _t(a, dev)
    col = A[:, 3]                          # stride 7
    assert col.stride(0) == 7
    vl = torch.empty(100, dtype=torch.float64, device=dev)[::2]
    vr = torch.empty(100, dtype=torch.float64, device=dev)[::2]
    hh.Weno(50, 3, 1e-6, arith="exact").reconstruct(col, vl, vr)
    ol, orr = oracle.reconstruct(a[:, 3], 3, 1e-6)
    assert np.array_equal(_h(vl), ol) and np.array_equal(_h(vr), orr)


def test_reconstruct_lines_both_directions(oracle, hh, dev):
    import torch
    n1, n2 = 70, 45
    rng = np.random.default_rng(9)
    u = rng.standard_normal(n1 * n2)
    U = _t(u, dev)
    vl = torch.empty(n1 * n2, dtype=torch.float64, device=dev)
    vr = torch.empty_like(vl)
    w1 = hh.Weno(n1, 3, 1e-6, arith="exact")
    w1.reconstruct_lines(U, n2, 1, n1, vl, vr, 1, n1)      # x1 lines (contiguous)
    L, R = _h(vl), _h(vr)
    for j in range(n2):
        ol, orr = oracle.reconstruct(u[j * n1:(j + 1) * n1], 3, 1e-6)
        assert np.array_equal(L[j * n1:(j + 1) * n1], ol) and np.array_equal(R[j * n1:(j + 1) * n1], orr)
    w2 = hh.Weno(n2, 3, 1e-6, arith="exact")
    w2.reconstruct_lines(U, n1, n1, 1, vl, vr, n1, 1)      # x2 lines (stride n1)
    L, R = _h(vl), _h(vr)
    for i in range(n1):
        ol, orr = oracle.reconstruct(u[i::n1], 3, 1e-6)
        assert np.array_equal(L[i::n1], ol) and np.array_equal(R[i::n1], orr)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_nonuniform_cnu_and_reconstruct(oracle, hh, dev, k):
    rng = np.random.default_rng(300 + k)
    n = 97
    xe = np.cumsum(rng.uniform(0.05, 1.0, n + 1))
    w = hh.Weno(n, k, 1e-6, xedges=xe)
    assert np.array_equal(w.cnu, oracle.calc_cnu(k, xe))
    v = rng.standard_normal(n)
    vl, vr = w.reconstruct(_t(v, dev))
    ol, orr = oracle.reconstruct(v, k, 1e-6, cnu=oracle.calc_cnu(k, xe))
    assert np.array_equal(_h(vl), ol) and np.array_equal(_h(vr), orr)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_reference_calc_cnu_and_nonuniform_pulse(hh, dev, k):
    """test/test_hrweno.f90:69-161 on the device."""
    from hrweno_hip import tables
    nc = 30
    xe = 3.0 * np.arange(nc + 1) / nc
    cnu = hh.Weno(nc, k, 1e-6, xedges=xe).cnu
    _, c = tables(k)
    assert np.allclose(cnu, np.broadcast_to(c.T, cnu.shape), rtol=1e-5, atol=0)
    xe = (np.arange(nc + 1) / nc) ** 3
    v = np.zeros(nc)
    v[9:20] = 1.0
    vl, vr = hh.Weno(nc, k, xedges=xe).reconstruct(_t(v, dev))
    assert np.max(np.abs(_h(vl) - v)) < 1e-6 and np.max(np.abs(_h(vr) - v)) < 1e-6


def test_reconstruct_host_arrays(oracle, hh, dev):
    rng = np.random.default_rng(11)
    v = rng.standard_normal(333)
    vl, vr = hh.Weno(333, 3, 1e-6, arith="exact").reconstruct_host(v)
    ol, orr = oracle.reconstruct(v, 3, 1e-6)
    assert np.array_equal(vl, ol) and np.array_equal(vr, orr)
    big = rng.standard_normal((333, 3))
    vl, vr = hh.Weno(333, 3, 1e-6, arith="exact").reconstruct_host(big[:, 1])
    ol, orr = oracle.reconstruct(big[:, 1].copy(), 3, 1e-6)
    assert np.array_equal(vl, ol) and np.array_equal(vr, orr)


@pytest.mark.parametrize("bad,msg", [
    (dict(ncells=0), "Invalid input 'ncells'. Valid range: ncells > 0."),
    (dict(k=4), "Invalid input 'k'. Valid range: 1 <= k <= 3."),
    (dict(eps=1e-17), "Invalid input 'eps'. Valid range: eps > epsilon."),
])
def test_weno_constructor_errors(hh, dev, bad, msg):
    args = dict(ncells=10, k=3, eps=1e-6)
    args.update(bad)
    with pytest.raises(hh.HrwError) as ei:
        hh.Weno(**args)
    assert str(ei.value) == msg


# ------------------------------------------------------------------------------ fluxes
def test_fluxes_known_answers(hh, dev):
    """test/test_fluxes.f90:27-71: f = u*x*t with x = 3, t = 5 -> (0 + t*x)*u."""
    vm = _t([2.0, 2.0, -2.0], dev)
    vp = _t([2.0, -4.0, 4.0], dev)
    x = _t([3.0, 3.0, 3.0], dev)
    href = np.array([30.0, 30.0, -30.0])
    g = hh.godunov(vm, vp, x, 5.0, fn=hh.FN_LINEAR, coef=(0.0, 5.0))
    lf = hh.lax_friedrichs(vm, vp, x, 5.0, alpha=15.0, fn=hh.FN_LINEAR, coef=(0.0, 5.0))
    assert np.allclose(_h(g), href, rtol=1e-8, atol=0)
    assert np.allclose(_h(lf), href, rtol=1e-8, atol=0)


def test_burgers_fluxes_bitwise(oracle, hh, dev):
    rng = np.random.default_rng(12)
    vm, vp = rng.standard_normal(5000), rng.standard_normal(5000)
    vp[:300] = vm[:300]
    g = _h(hh.godunov(_t(vm, dev), _t(vp, dev)))
    lf = _h(hh.lax_friedrichs(_t(vm, dev), _t(vp, dev), alpha=1.7))
    for i in range(0, 5000, 7):
        fm, fp = vm[i] * vm[i] / 2, vp[i] * vp[i] / 2   # v*v, not numpy's pow-based v**2
        assert g[i] == oracle.godunov(fm, fp, vm[i], vp[i])
        assert lf[i] == oracle.lax_friedrichs(fm, fp, vm[i], vp[i], 1.7)
