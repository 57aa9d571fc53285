"""Seeded synthetic inputs for the benchmark configurations (host numpy, FP64).

Random numbers come from SplitMix64: draw i of stream `seed` is mix(seed + (i+1)*G) with
G = 0x9E3779B97F4A7C15; U(0,1) = (x >> 11) * 2**-53; N(0,1) by Box-Muller.  Initial data are
sampled at cell centres (example/example1_burgers_1d_fv.f90:50).  The same arrays feed the
device path and the CPU checker.
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_MASK = (1 << 64) - 1


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Draws start .. start+count-1 of the SplitMix64 stream seeded with `seed`."""
    i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & _MASK) + i * _G
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


class Rng:
    def __init__(self, seed: int):
        self.seed = seed & _MASK
        self.pos = 0

    def uniform(self, count: int = 1, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        z = splitmix64(self.seed, count, self.pos)
        self.pos += count
        u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, count: int = 1, mean: float = 0.0, std: float = 1.0) -> np.ndarray:
        u1 = 1.0 - self.uniform(count)          # (0, 1]
        u2 = self.uniform(count)
        return mean + std * np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)

    def integers(self, count: int, lo: int, hi: int) -> np.ndarray:
        """Uniform integers in [lo, hi]."""
        return (lo + np.floor(self.uniform(count) * (hi - lo + 1))).astype(np.int64).clip(lo, hi)


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def linear_centres(xmin: float, xmax: float, n: int):
    """Edges, centres and widths of grid1%linear (src/hrweno_grids.f90:76-78, 246-247)."""
    rx = (xmax - xmin) / n
    e = xmin + rx * np.arange(n + 1, dtype=np.float64)
    return e, (e[:-1] + e[1:]) / 2, e[1:] - e[:-1]


@dataclass
class Burgers1DCase:
    """One periodic Burgers line: u0, grid, k, dt, steps or tout."""
    n: int
    u0: np.ndarray
    xmin: float = -1.0
    xmax: float = 1.0
    k: int = 3
    eps: float = 1e-6
    dt: float = 0.0
    tout: float = 0.0
    steps: int = 0
    width: np.ndarray = field(default=None, repr=False)
    seed: int = 0


def cfl_dt(u0: np.ndarray, dx: float, cfl: float = 0.4) -> float:
    return cfl * dx / float(np.max(np.abs(u0)))


def smooth_step_ic(x: np.ndarray, rng: Rng) -> np.ndarray:
    """0.5 + sin(pi x) + h*H(x - xs), h ~ U(0,1), xs ~ U(-0.5, 0.5)."""
    h = rng.uniform(1)[0]
    xs = rng.uniform(1, -0.5, 0.5)[0]
    return 0.5 + np.sin(np.pi * x) + h * (x >= xs)


def config1(n: int = 1000, seed: int = 1) -> Burgers1DCase:
    """Burgers, 1000 cells on [-1,1], periodic, WENO5-LF-RK3, CFL 0.4, to t = 0.5."""
    e, xc, w = linear_centres(-1.0, 1.0, n)
    u0 = smooth_step_ic(xc, Rng(seed))
    return Burgers1DCase(n=n, u0=u0, dt=cfl_dt(u0, 2.0 / n), tout=0.5, width=w, seed=seed)


def config2(n: int = 1 << 24, seed: int = 2, steps: int = 100) -> Burgers1DCase:
    """Burgers, 2**24 cells: 8 seeded sine modes + 16 seeded jumps, 100 RK3 steps."""
    rng = Rng(seed)
    e, xc, w = linear_centres(-1.0, 1.0, n)
    amp = rng.uniform(8, 0.0, 0.5)
    kap = rng.integers(8, 1, 64)
    phi = rng.uniform(8, 0.0, 2 * np.pi)
    xj = rng.uniform(16, -1.0, 1.0)
    hj = rng.normal(16, 0.0, 0.5)
    u0 = np.zeros(n)
    for m in range(8):
        u0 += amp[m] * np.sin(np.pi * kap[m] * xc + phi[m])
    for j in range(16):
        u0 += hj[j] * (xc >= xj[j])
    return Burgers1DCase(n=n, u0=u0, dt=cfl_dt(u0, 2.0 / n), steps=steps, width=w, seed=seed)


def config3(nlines: int = 4096, seed: int = 3, steps: int = 200):
    """4096 independent periodic Burgers lines, lengths log-uniform in [256, 65536],
    k in {2, 3} (50/50), config-1 style data with own seed, own CFL dt."""
    rng = Rng(seed)
    lens = np.rint(np.exp(rng.uniform(nlines, np.log(256.0), np.log(65536.0)))).astype(np.int64)
    ks = np.where(rng.uniform(nlines) < 0.5, 2, 3)
    cases = []
    for l in range(nlines):
        n = int(lens[l])
        e, xc, w = linear_centres(-1.0, 1.0, n)
        u0 = smooth_step_ic(xc, Rng((seed << 32) + l))
        cases.append(Burgers1DCase(n=n, u0=u0, k=int(ks[l]), dt=cfl_dt(u0, 2.0 / n),
                                   steps=steps, width=w, seed=(seed << 32) + l))
    return cases


def config5(n: int = 1 << 20, seed: int = 5, noise_seed: int = 55, steps: int = 400):
    """Adjoint form: config-1 style Burgers data on 2**20 periodic cells, 400 RK3 steps.  The
    loss is 0.5*|u_N - target|^2 with target = the forward run from u0 + N(0, 0.01) per cell;
    returns (case, noise)."""
    e, xc, w = linear_centres(-1.0, 1.0, n)
    u0 = smooth_step_ic(xc, Rng(seed))
    noise = Rng(noise_seed).normal(n, 0.0, 0.01)
    return Burgers1DCase(n=n, u0=u0, dt=cfl_dt(u0, 2.0 / n), steps=steps, width=w,
                         seed=seed), noise


def config5_batch(nlines: int = 256, seed: int = 5, steps: int = 400, lo: int = 1024,
                  hi: int = 16384):
    """Batched adjoint form: 256 periodic lines, lengths log-uniform in [1024, 16384], WENO5,
    config-1 style data with own seed and CFL dt; each line's target from its own N(0, 0.01)
    perturbation.  Returns (cases, noises)."""
    rng = Rng(seed)
    lens = np.rint(np.exp(rng.uniform(nlines, np.log(float(lo)), np.log(float(hi))))).astype(np.int64)
    cases, noises = [], []
    for l in range(nlines):
        n = int(lens[l])
        e, xc, w = linear_centres(-1.0, 1.0, n)
        u0 = smooth_step_ic(xc, Rng((seed << 32) + l))
        cases.append(Burgers1DCase(n=n, u0=u0, k=3, dt=cfl_dt(u0, 2.0 / n), steps=steps,
                                   width=w, seed=(seed << 32) + l))
        noises.append(Rng((55 << 32) + l).normal(n, 0.0, 0.01))
    return cases, noises


@dataclass
class PBE2DCase:
    n1: int
    n2: int
    u0: np.ndarray           # u((j-1)*n1 + i)
    dt: float
    steps: int
    g1: tuple = (1.0, 0.5)
    g2: tuple = (0.5, 0.0)
    k: int = 3
    eps: float = 1e-6


def config4(n: int = 2048, seed: int = 4, steps: int = 500) -> PBE2DCase:
    """2D PBE on the unit square, n x n, Gaussian pulse, centre ~ U(0.2,0.4)^2, width 0.05,
    g1 = 1 + 0.5 x1, g2 = 0.5, mstvd, dt = 0.4 / (max g1/dx1 + max g2/dx2)."""
    rng = Rng(seed)
    c = rng.uniform(2, 0.2, 0.4)
    e, xc, w = linear_centres(0.0, 1.0, n)
    X1, X2 = np.meshgrid(xc, xc, indexing="xy")   # X1[j, i] = x1_i, X2[j, i] = x2_j
    u0 = np.exp(-((X1 - c[0]) ** 2 + (X2 - c[1]) ** 2) / (2 * 0.05 ** 2)).ravel()
    dx = 1.0 / n
    dt = 0.4 / (1.5 / dx + 0.5 / dx)
    return PBE2DCase(n1=n, n2=n, u0=u0, dt=dt, steps=steps)


def steps_tout(t0: float, dt: float, steps: int) -> float:
    """A tout at which the reference loop control (strict is_done) takes exactly `steps` steps
    from t0: the FP64 time after `steps-1` steps plus dt/2."""
    t = t0
    for _ in range(steps - 1):
        t = t + dt
    return t + dt / 2
