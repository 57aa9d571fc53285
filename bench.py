#!/usr/bin/env python3
"""Benchmark of the fused WENO5 + Lax-Friedrichs + SSP-RK3 path (FP64) on one MI355X.

Workload: 1D Burgers on 2**24 periodic cells of [-1, 1] (8 seeded sine modes + 16 seeded
jumps), k = 3, eps = 1e-6, alpha = max|u| of each stage input, fixed dt at CFL 0.4.  A "step"
is one RK3 time step of the whole line: three fused stage launches.  The state stays resident
in HBM; only the integrate calls are timed (HIP events on the launch stream).

Prints ONE JSON line.  `roofline` prices the stage kernel against HBM (64 algorithmic bytes
per cell-step = 16 + 24 + 24 over the three stages); `cpu_baseline` times the reference's own
compiled hrweno_weno/hrweno_fluxes (oracle/_ref) on a bounded sample on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "hr-weno_amd"))

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_STEP = 64.0          # RK3, one fused pass per stage: 16 + 24 + 24
LAUNCHES_PER_STEP = 3
METRIC = "FP64 cell-updates/s & % HBM peak: plain vs batched vs adjoint, ≤1e-12 rel"
TRAFFIC_FILE = os.path.join(REPO, "profiles", "stage1d_burgers_2p24_traffic.json")


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(seconds: float):
    """The reference's compiled reconstruct + Lax-Friedrichs (oracle/_ref) driving periodic
    Burgers RK3 on a 2**20-cell sample of the same workload, one core."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    from hrweno_hip import workloads as W
    import oracle as O
    n = 1 << 20
    c = W.config2(n=n)
    if O.Reference.available():
        r = O.Reference()
        run = lambda steps: r.burgers_rk3(c.u0, c.width, steps, c.dt)  # noqa: E731
        kind = "reference"
        what = ("oracle/_ref: the reference's own hrweno_weno%reconstruct and lax_friedrichs "
                "(amdflang -O3, no FMA) in an example-1-style periodic RHS with the RK3 stage "
                "lines of src/hrweno_tvdode.f90:162-167")
    else:
        o = O.Oracle()
        o.set_threads(1)
        run = lambda steps: o.burgers_rk(c.u0, c.width, tout=W.steps_tout(0.0, c.dt, steps),  # noqa: E731
                                         dt=c.dt)
        kind = "port"
        what = "oracle C restatement (gcc -O3, no FMA)"
    # The reference keeps O(n) automatic arrays on the stack (src/hrweno_weno.f90:163), so the
    # timed calls run on a thread with a 1 GiB stack.
    import threading
    res = {}

    def worker():
        t0 = time.perf_counter()
        run(2)
        per_step = (time.perf_counter() - t0) / 2
        res["steps"] = steps = max(2, int(seconds / per_step))
        t0 = time.perf_counter()
        run(steps)
        res["time"] = time.perf_counter() - t0

    old_size = threading.stack_size(1 << 30)
    th = threading.Thread(target=worker)
    th.start()
    th.join()
    threading.stack_size(old_size)
    steps, dt = res["steps"], res["time"]
    return {"value": n * steps / dt, "unit": "cell-updates/s", "cores": 1, "kind": kind,
            "sample": f"{n} cells x {steps} RK3 steps of the same workload (WENO5-LF, periodic), "
                      f"{dt:.1f} s on one host core; {what}"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--arith", choices=["fast", "exact"], default="fast")
    ap.add_argument("--cells", type=int, default=1 << 24)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--variant", type=int, default=0,
                    help="1D FAST kernel: 0 register-blocked, 1 LDS tile")
    ap.add_argument("--cells-per-lane", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain device copy of --cells doubles (20 times), the "
                         "known-size kernel that calibrates the FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path shards "
              "by domain decomposition, which this build does not launch.", file=sys.stderr)
        return 2

    import torch
    import hrweno_hip as hh
    from hrweno_hip import workloads as W

    dev = torch.device("cuda:0")
    if args.copy_probe:
        src = torch.rand(args.cells, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(20):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: 20 copies of {args.cells * 8} bytes")
        return 0
    log(f"generating workload: {args.cells} cells")
    c = W.config2(n=args.cells)
    prob = hh.Burgers1D(c.n, 3, 1e-6, flux=hh.LAX_FRIEDRICHS, bc=hh.BC_PERIODIC, alpha=-1.0,
                        xmin=-1.0, xmax=1.0, arith=args.arith)
    prob.tune(args.variant, args.cells_per_lane)
    ode = hh.RkTvd(prob, order=3)
    u = torch.as_tensor(c.u0, device=dev).clone()
    t = 0.0
    if args.warmup > 0:
        t = ode.integrate(u, t, W.steps_tout(t, c.dt, args.warmup), c.dt)
    torch.cuda.synchronize()
    log(f"warmup done ({args.warmup} steps)")

    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    fe0 = ode.fevals
    tout = W.steps_tout(t, c.dt, args.steps)
    torch.cuda.synchronize()
    e0.record(stream)
    t = ode.integrate(u, t, tout, c.dt)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    steps_done = (ode.fevals - fe0) // 3
    assert steps_done == args.steps, (steps_done, args.steps)
    if not bool(torch.isfinite(u).all()):
        raise RuntimeError("non-finite state after the timed steps")
    log(f"timed {args.steps} steps: {ms:.3f} ms")

    value = c.n * args.steps / (ms / 1e3)
    launch_ms = ms / (LAUNCHES_PER_STEP * args.steps)
    bytes_per_launch = BYTES_PER_CELL_STEP / LAUNCHES_PER_STEP * c.n
    achieved = bytes_per_launch / (launch_ms / 1e3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": (("hrw::stage1d_fast_kernel<3,LF,periodic,alphamax,%d>" % args.cells_per_lane)
                           if args.arith == "fast" and args.variant == 0 else
                           "hrw::stage1d_kernel<3,LF,periodic,%s,alphamax,4>" % args.arith),
                "avg_launch_us": round(launch_ms * 1e3, 2),
                "algorithmic_bytes_per_launch": bytes_per_launch}

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(args.cpu_seconds)

    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded SplitMix64: 8 sine modes + 16 jumps)",
        "config": {"workload": "burgers1d_weno5_lf_rk3_periodic_2^24cells", "cells": c.n,
                   "k": 3, "eps": 1e-6, "flux": "lax_friedrichs, alpha=max|u| per stage",
                   "integrator": "rktvd order 3", "cfl": 0.4, "dt": c.dt,
                   "arith": args.arith, "variant": args.variant,
                   "cells_per_lane": args.cells_per_lane, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
