#!/usr/bin/env python3
"""Benchmarks of the fused WENO -> flux -> divergence -> TVD-stage path (FP64) on one MI355X.

Default workload (the headline line): 1D Burgers on 2**24 periodic cells of [-1, 1] (8 seeded
sine modes + 16 seeded jumps), WENO5 (k = 3, eps = 1e-6), Lax-Friedrichs with alpha = max|u|
of each stage input, SSP-RK3 at CFL 0.4.  A "step" is one RK3 time step of the whole line:
three fused stage launches.  The state stays resident in HBM; only the integrate calls are
timed (HIP events on the launch stream).

The other legs of the metric run in the same invocation, each at its own configuration and
step count (whatever --steps says for the headline), and are reported as sub-objects of the one
JSON line, each with its own `roofline` from this run's own timing:

  `batch`         -- 4096 independent periodic lines of log-uniform length in [256, 65536],
                     k in {2, 3}, own dt, 200 RK3 steps, one launch per stage for all lines
                     (autotuned, graph replayed) vs one launch per line (20 steps);
  `adjoint`       -- gradient of 0.5*|u_N - target|^2 w.r.t. u0, 2^20 cells x 400 RK3 steps:
                     the automatic tape (every stage input kept) and a checkpoint every 20 steps,
                     each as gradient time / plain forward time, and the reverse kernel's roofline;
  `adjoint_batch` -- the same over 256 lines of 1024..16384 cells vs one adjoint per line;
  `adjoint_general` -- example 1's composition (Godunov, zero-gradient ends) at 2^20 cells x 400
                     steps through the general reverse pass: the rktvd gradient and the recorded
                     mstvd integrator's, each over its plain forward;
  `pbe`           -- 2D population balance on 2048 x 2048, 5-step multistep method, 500 steps
                     (the EXACT form: the multistep method amplifies rounding, so only the
                     reference's own arithmetic meets 1e-12 over 500 steps).

`--workload X` runs one leg alone as the headline (`--legs none` drops the sub-objects).

`roofline` prices the dominant kernel against HBM with the algorithmic bytes per cell-step
(RK3 1D: 16 + 24 + 24 = 64; multistep 2D: 40; reverse RK3 stage: 88); `traffic` is the HBM
bytes per launch measured by rocprofv3 FETCH_SIZE/WRITE_SIZE passes (profiles/traffic.json).
`cpu_baseline` times the reference's own compiled hrweno_weno/hrweno_fluxes (oracle/_ref) on a
bounded sample on one core, and beside it (`all_cores`, labelled apart) the oracle's C
restatement with OpenMP on every core the job may use, with the host's CPU model and core
counts (`host`).
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
METRIC = "FP64 cell-updates/s & % HBM peak: plain vs batched vs adjoint, ≤1e-12 rel"
TRAFFIC_FILE = os.path.join(REPO, "profiles", "traffic.json")


# steps of each workload when it is not the headline (the configurations' own horizons)
LEG_STEPS = {"plain": 100, "batch": 200, "pbe": 500, "adjoint": 400, "adjoint_batch": 400,
             "adjoint_general": 400}


def release(torch) -> None:
    """Return the device memory of the leg just run before the next one allocates."""
    import gc
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(seconds: float):
    """The reference's compiled reconstruct + Lax-Friedrichs (oracle/_ref) driving periodic
    Burgers RK3 on a 2**20-cell sample of the headline workload, one core."""
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
    out = {"value": n * steps / dt, "unit": "cell-updates/s", "cores": 1, "kind": kind,
           "sample": f"{n} cells x {steps} RK3 steps of the same workload (WENO5-LF, periodic), "
                     f"{dt:.1f} s on one host core; {what}"}
    out["host"] = host_info()
    out["all_cores"] = cpu_all_cores(seconds)
    return out


def host_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def cpu_all_cores(seconds: float):
    """The oracle's C restatement (OpenMP over cells, the reference's own parallel pattern of
    example2:95-114 applied to one line) on every host core this job may use, on a 2**22-cell
    sample of the headline workload; labelled apart from the one-core reference figure."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    from hrweno_hip import workloads as W
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0))
    n = 1 << 22
    c = W.config2(n=n)
    o = O.Oracle()
    o.set_threads(threads)
    run = lambda steps: o.burgers_rk(c.u0, c.width, tout=W.steps_tout(0.0, c.dt, steps),  # noqa: E731
                                     dt=c.dt)
    t0 = time.perf_counter()
    run(2)
    steps = max(2, int(seconds / ((time.perf_counter() - t0) / 2)))
    t0 = time.perf_counter()
    run(steps)
    dt = time.perf_counter() - t0
    return {"value": n * steps / dt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} cells x {steps} RK3 steps, {dt:.1f} s on {threads} threads; oracle C "
                      "restatement (gcc -O3, no FMA, OpenMP)"}


def traffic_for(kernel_key: str):
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            return json.load(f).get(kernel_key, {}).get("hbm_bytes_per_launch")
    return None


def timed(torch, fn):
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1)


def roofline(bytes_per_launch, launch_ms, kernel, key):
    achieved = bytes_per_launch / (launch_ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_for(key),
            "kernel": kernel, "avg_launch_us": round(launch_ms * 1e3, 2),
            "algorithmic_bytes_per_launch": bytes_per_launch}


def run_plain(args, torch, hh, W, dev):
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
    fe0 = ode.fevals
    tout = W.steps_tout(t, c.dt, args.steps)
    res = {}
    ms = timed(torch, lambda: res.setdefault("t", ode.integrate(u, t, tout, c.dt)))
    assert (ode.fevals - fe0) // 3 == args.steps
    if not bool(torch.isfinite(u).all()):
        raise RuntimeError("non-finite state after the timed steps")
    log(f"timed {args.steps} steps: {ms:.3f} ms")
    # the library's automatic choice: lines of 2^21 cells or more run the persistent form with 4
    # cells per lane, 2^20..2^21 one chunk per wavefront with 6, shorter ones with 4
    automatic = args.variant == 0 and not args.cells_per_lane
    persistent = args.variant == 2 or (automatic and c.n >= (1 << 21))
    cpl = args.cells_per_lane or (6 if not persistent and c.n >= (1 << 20) else 4)
    if args.arith != "fast" or args.variant == 1:
        kernel = "hrw::stage1d_kernel<3,LF,periodic,%s,alphamax,4>" % args.arith
    else:
        kernel = ("hrw::%s<3,LF,periodic,alphamax,%d,NPW> (NPW = 0 for stage 1, 1 for stages 2-3)"
                  % ("stage1d_pf_kernel" if persistent else "stage1d_fast_kernel", cpl))
    rl = roofline(64.0 / 3 * c.n, ms / (3 * args.steps), kernel, "stage1d_2p24")
    config = {"workload": "burgers1d_weno5_lf_rk3_periodic_2^24cells", "cells": c.n, "k": 3,
              "eps": 1e-6, "flux": "lax_friedrichs, alpha=max|u| per stage",
              "integrator": "rktvd order 3", "cfl": 0.4, "dt": c.dt, "arith": args.arith,
              "variant": args.variant, "cells_per_lane": cpl,
              "form": "persistent" if persistent else "one chunk per wavefront",
              "parallelism": "single GPU"}
    return c.n * args.steps / (ms / 1e3), ms / args.steps, rl, config, \
        "synthetic (seeded SplitMix64: 8 sine modes + 16 jumps)"


def run_batch(args, torch, hh, W, dev):
    import numpy as np
    steps = args.steps
    log("generating workload: 4096 lines")
    cases = W.config3(steps=steps)
    lens = np.array([c.n for c in cases])
    ks = np.array([c.k for c in cases])
    dts = np.array([c.dt for c in cases])
    total = int(lens.sum())
    b = hh.BurgersBatch(lens, ks, dts)
    u0 = b.pack([c.u0 for c in cases], dev)
    u = u0.clone()
    r = b.tune(u)
    form = b.config()[1]
    b.step(u, args.warmup, graph=True)
    ms = timed(torch, lambda: b.step(u, steps, graph=True))
    log(f"batched (graph, {'persistent' if form else 'chunk'} form, {r} cells/lane): "
        f"{steps} steps {ms:.3f} ms")
    ms_nograph = timed(torch, lambda: b.step(u, steps, graph=False))
    log(f"batched (no graph): {steps} steps {ms_nograph:.3f} ms")
    # one-launch-per-line path on a bounded number of steps
    pl_steps = max(1, min(steps, args.per_line_steps))
    up = u0.clone()
    ms_pl = timed(torch, lambda: b.step_per_line(up, pl_steps))
    log(f"per-line: {pl_steps} steps {ms_pl:.3f} ms")
    value = total * steps / (ms / 1e3)
    per_line_value = total * pl_steps / (ms_pl / 1e3)
    kname = "hrw::batch_pf_kernel" if form else "hrw::batch_stage_kernel"
    rl = roofline(64.0 / 3 * total, ms / (3 * steps), "%s<LF,alphamax,%d>" % (kname, r), "batch_c3")
    config = {"workload": "burgers1d_batch_4096_lines_log_uniform_256_65536", "lines": len(cases),
              "cells": total, "k": "2 or 3 per line", "integrator": "rktvd order 3",
              "cells_per_lane": r, "form": "persistent" if form else "chunk", "graph": True,
              "parallelism": "single GPU",
              "batched_no_graph_cell_updates_per_s": total * steps / (ms_nograph / 1e3),
              "per_line_cell_updates_per_s": per_line_value, "per_line_steps": pl_steps,
              "speedup_vs_per_line": value / per_line_value}
    return value, ms / steps, rl, config, "synthetic (seeded SplitMix64, config-1 style lines)"


def run_pbe(args, torch, hh, W, dev):
    """The multistep method amplifies rounding on this problem (a 1-ulp change of u0 reaches
    ~1e-1 by step 500 in the reference's own arithmetic), so over its 500 steps only the EXACT
    form (the reference's evaluation order, bit-identical) meets 1e-12; that is the timed one.
    The FAST form (within 1e-12 for ~200 steps) is timed beside it for comparison."""
    c = W.config4(steps=args.steps)
    ncell = c.n1 * c.n2
    arith = "exact" if args.pbe_arith == "auto" else args.pbe_arith
    out = {}
    for form in dict.fromkeys([arith, "fast"]):
        ode = hh.MsTvd(hh.PBE2D(c.n1, c.n2, k=3, bc=hh.BC_INFLOW0_OUTFLOW, arith=form))
        u = torch.as_tensor(c.u0, device=dev).clone()
        t = ode.integrate(u, 0.0, W.steps_tout(0.0, c.dt, 4 + args.warmup), c.dt)   # startup + warmup
        tout = W.steps_tout(t, c.dt, args.steps)
        fe0 = ode.fevals
        ms = timed(torch, lambda: ode.integrate(u, t, tout, c.dt))
        assert ode.fevals - fe0 == args.steps
        out[form] = ms
        log(f"pbe {form}: {args.steps} steps {ms:.3f} ms")
    ms = out[arith]
    rl = roofline(40.0 * ncell, ms / args.steps,
                  "hrw::pbe2d_kernel<3,inflow0_outflow,%s>" % arith, "pbe2d_2048_" + arith)
    config = {"workload": "pbe2d_2048x2048_weno5_godunov_mstvd", "cells": ncell,
              "integrator": "mstvd (5-step, 3rd order)", "g1": "1+0.5*x1", "g2": "0.5",
              "bc": "zero inflow, upwind outflow", "arith": arith,
              "parity": "exact: bit-identical to the reference arithmetic over all 500 steps"
                        if arith == "exact" else "fast: within 1e-12 over ~20 steps only (the "
                                                 "multistep method amplifies rounding)",
              "fast_cell_updates_per_s": ncell * args.steps / (out["fast"] / 1e3),
              "fast_roofline_frac": roofline(40.0 * ncell, out["fast"] / args.steps, "", "")["frac"],
              "parallelism": "single GPU"}
    return ncell * args.steps / (ms / 1e3), ms / args.steps, rl, config, \
        "synthetic (seeded Gaussian pulse)"


def run_adjoint(args, torch, hh, W, dev):
    """Gradient of 0.5*|u_N - target|^2 w.r.t. u0 for 2**20 cells x 400 RK3 steps: the
    recording forward + the backward.  The default checkpoint interval is automatic (the whole
    tape fits in HBM, so the backward is the reverse kernels alone); the low-memory schedule
    (a checkpoint every 20 steps, segments recomputed) is timed beside it."""
    c, noise = W.config5(steps=args.steps)
    n, steps = c.n, c.steps
    prob = hh.Burgers1D(n)
    adj = hh.Adjoint(prob, nsteps=steps, dt=c.dt, ckpt_every=args.ckpt_every)
    u0 = torch.as_tensor(c.u0, device=dev).clone()
    target = adj.forward(torch.as_tensor(c.u0 + noise, device=dev)).clone()
    res = {}
    adj.backward(adj.forward(u0) - target)          # warm-up
    ms_f = timed(torch, lambda: res.__setitem__("uN", adj.forward(u0)))
    gN = res["uN"] - target
    ms_b = timed(torch, lambda: res.__setitem__("g0", adj.backward(gN)))
    if not bool(torch.isfinite(res["g0"]).all()):
        raise RuntimeError("non-finite gradient")
    # the plain forward of the same steps (no recording)
    ode = hh.RkTvd(prob, order=3)
    up = u0.clone()
    ms_p = timed(torch, lambda: ode.integrate(up, 0.0, W.steps_tout(0.0, c.dt, steps), c.dt))
    # the reverse kernels alone: every stage input kept by the forward, nothing recomputed
    full = hh.Adjoint(prob, nsteps=steps, dt=c.dt, ckpt_every=steps)
    full.forward(u0)
    full.backward(gN)
    ms_r = timed(torch, lambda: full.backward(gN))
    del full
    # the low-memory schedule: a checkpoint every 20 steps, each segment recomputed once
    low = hh.Adjoint(prob, nsteps=steps, dt=c.dt, ckpt_every=min(20, steps))
    low.backward(low.forward(u0) - target)
    ms_lf = timed(torch, lambda: low.forward(u0))
    ms_lb = timed(torch, lambda: low.backward(gN))
    info, linfo = adj.info(), low.info()
    del low
    log(f"adjoint: forward(recording) {ms_f:.2f} ms, backward {ms_b:.2f} ms, plain forward "
        f"{ms_p:.2f} ms, reverse-only {ms_r:.2f} ms; checkpoint every 20: forward "
        f"{ms_lf:.2f} ms, backward {ms_lb:.2f} ms")
    rl = roofline(88.0 / 3 * n, ms_r / (3 * steps), "hrw::adjoint_stage_kernel<2,alphamax>",
                  "adjoint_2p20")
    config = {"workload": "burgers1d_weno5_lf_rk3_adjoint_2^20cells_400steps", "cells": n,
              "steps": steps, "ckpt_every": info["ckpt_every"], "k": 3, "eps": 1e-6,
              "loss": "0.5*|u_N - F(u0 + N(0,0.01))|^2", "parallelism": "single GPU",
              "value_is": "cell-steps of the whole gradient (recording forward + backward) per s",
              "forward_only_cell_updates_per_s": n * steps / (ms_p / 1e3),
              "gradient_over_forward": (ms_f + ms_b) / ms_p,
              "backward_over_forward": ms_b / ms_p,
              "reverse_only_over_forward": ms_r / ms_p,
              "device_bytes": info["device_bytes"],
              "ckpt20_gradient_over_forward": (ms_lf + ms_lb) / ms_p,
              "ckpt20_device_bytes": linfo["device_bytes"]}
    return n * steps / ((ms_f + ms_b) / 1e3), (ms_f + ms_b) / steps, rl, config, \
        "synthetic (seeded SplitMix64, config-1 style data; target from N(0,0.01) noise)"


def run_adjoint_general(args, torch, hh, W, dev):
    """Example 1's own composition (the endpoint Godunov flux, zero-gradient ends, WENO5,
    example/example1_burgers_1d_fv.f90:72-109) at 2**20 cells x 400 steps through the general
    reverse pass: the rktvd gradient of 0.5*|u_N - target|^2 (every stage input kept, so the
    backward is the reverse launches alone) and the recorded mstvd integrator's (hrw_ode_tape /
    hrw_ode_adjoint, the four RK3 starting steps included), each against its plain forward."""
    import numpy as np
    c, _ = W.config5(steps=args.steps)
    n, steps, dt = c.n, c.steps, c.dt
    prob = hh.Burgers1D(n, flux=hh.GODUNOV, bc=hh.BC_ZERO_GRADIENT)
    prob.tune(args.variant, args.cells_per_lane)     # variant 1: the LDS-tile kernels both ways
    adj = hh.Adjoint(prob, nsteps=steps, dt=dt, ckpt_every=steps)
    u0 = torch.as_tensor(c.u0, device=dev).clone()
    # the target from a smooth perturbation: with per-cell noise the endpoint Godunov flux
    # turns unstable at the transonic point of this state within 400 steps (the reference's
    # arithmetic too; the oracle shows it at 2^18 cells)
    _, xc, _ = W.linear_centres(-1.0, 1.0, n)
    pert = 0.01 * np.sin(3.0 * np.pi * xc)
    target = adj.forward(torch.as_tensor(c.u0 + pert, device=dev)).clone()
    res = {}
    adj.backward(adj.forward(u0) - target)          # warm-up
    ms_f = timed(torch, lambda: res.__setitem__("uN", adj.forward(u0)))
    gN = res["uN"] - target
    ms_b = timed(torch, lambda: res.__setitem__("g0", adj.backward(gN)))
    if not bool(torch.isfinite(res["g0"]).all()):
        raise RuntimeError("non-finite gradient")
    tout = W.steps_tout(0.0, dt, steps)
    ode = hh.RkTvd(prob, order=3)
    up = u0.clone()
    ms_p = timed(torch, lambda: ode.integrate(up, 0.0, tout, dt))
    del adj

    # mstvd: a recording starts with the method (istate 1); loading the state saved before the
    # first step starts it again, and the recording reuses the storage of the previous one
    om = hh.MsTvd(prob).tape(1)
    um = u0.clone()
    s0 = om.save(um, 0.0)
    om.integrate(um, 0.0, tout, dt)                  # warm-up (allocates the tape)
    om.load(s0, um)
    ms_mf = timed(torch, lambda: om.integrate(um, 0.0, tout, dt))
    lam = um - target
    ms_mb = timed(torch, lambda: om.adjoint(lam))
    if not bool(torch.isfinite(lam).all()):
        raise RuntimeError("non-finite gradient")
    del om
    op = hh.MsTvd(prob)
    up = u0.clone()
    p0 = op.save(up, 0.0)
    op.integrate(up, 0.0, tout, dt)                  # warm-up
    op.load(p0, up)
    ms_mp = timed(torch, lambda: op.integrate(up, 0.0, tout, dt))
    log(f"adjoint (example 1 composition): rktvd forward(recording) {ms_f:.2f} ms, backward "
        f"{ms_b:.2f} ms, plain forward {ms_p:.2f} ms; mstvd recording {ms_mf:.2f} ms, adjoint "
        f"{ms_mb:.2f} ms, plain forward {ms_mp:.2f} ms")
    tile = args.variant == 1
    rl = roofline(88.0 / 3 * n, ms_b / (3 * steps),
                  "hrw::rhs_vjp_kernel<3,GODUNOV,zero-gradient,fixed>" if tile
                  else "hrw::rhs_vjp5_kernel<GODUNOV,zero-gradient,fixed>",
                  "adjoint_general_2p20_tile" if tile else "adjoint_general_2p20")
    config = {"workload": "burgers1d_example1_godunov_zerograd_weno5_adjoint_2^20cells_400steps",
              "cells": n, "steps": steps, "k": 3, "eps": 1e-6,
              "loss": "0.5*|u_N - F(u0 + 0.01 sin(3 pi x))|^2", "parallelism": "single GPU",
              "value_is": "cell-steps of the rktvd gradient (recording forward + backward) per s",
              "gradient_over_forward": (ms_f + ms_b) / ms_p,
              "backward_over_forward": ms_b / ms_p,
              "mstvd_gradient_over_forward": (ms_mf + ms_mb) / ms_mp,
              "mstvd_recording_ms": ms_mf, "mstvd_adjoint_ms": ms_mb, "mstvd_forward_ms": ms_mp}
    return n * steps / ((ms_f + ms_b) / 1e3), (ms_f + ms_b) / steps, rl, config, \
        "synthetic (seeded SplitMix64, config-1 style data; target from a smooth perturbation)"


def run_adjoint_batch(args, torch, hh, W, dev):
    """The batched adjoint: 256 lines of 1024-16384 cells, every stage of every line in one
    launch, vs one adjoint object per line."""
    import numpy as np
    cases, noises = W.config5_batch(steps=args.steps)
    steps = args.steps
    lens = np.array([c.n for c in cases])
    b = hh.BurgersBatch(lens, np.full(lens.size, 3), np.array([c.dt for c in cases]))
    adj = hh.Adjoint(b, nsteps=steps, ckpt_every=args.ckpt_every)
    u0 = b.pack([c.u0 for c in cases], dev)
    target = adj.forward(b.pack([c.u0 + z for c, z in zip(cases, noises)], dev)).clone()
    res = {}
    adj.backward(adj.forward(u0) - target)
    ms = timed(torch, lambda: res.__setitem__("g", adj.backward(adj.forward(u0) - target)))
    total = int(lens.sum())
    pl_steps = max(1, min(steps, args.per_line_steps))
    per = [hh.Adjoint(hh.Burgers1D(int(c.n)), nsteps=pl_steps, dt=c.dt,
                      ckpt_every=args.ckpt_every) for c in cases]
    lines = [torch.as_tensor(c.u0, device=dev) for c in cases]

    def per_line():
        for a, x in zip(per, lines):
            a.backward(a.forward(x))
    per_line()
    ms_pl = timed(torch, per_line)
    value = total * steps / (ms / 1e3)
    pl_value = total * pl_steps / (ms_pl / 1e3)
    log(f"batched adjoint: {steps} steps {ms:.2f} ms; per-line {pl_steps} steps {ms_pl:.2f} ms")
    full = hh.Adjoint(b, nsteps=steps, ckpt_every=steps)   # reverse kernels alone
    gN = adj.forward(u0) - target
    full.forward(u0)
    full.backward(gN)
    ms_r = timed(torch, lambda: full.backward(gN))
    rl = roofline(88.0 / 3 * total, ms_r / (3 * steps), "hrw::adjoint_stage_kernel<2,alphamax>",
                  "adjoint_batch")
    config = {"workload": "burgers1d_batch_adjoint_256_lines_log_uniform_1024_16384",
              "lines": len(cases), "cells": total, "steps": steps, "k": 3,
              "ckpt_every": adj.info()["ckpt_every"], "parallelism": "single GPU",
              "value_is": "cell-steps of the whole gradient (recording forward + backward) per s",
              "per_line_cell_updates_per_s": pl_value, "per_line_steps": pl_steps,
              "speedup_vs_per_line": value / pl_value}
    return value, ms / steps, rl, config, "synthetic (seeded SplitMix64, config-1 style lines)"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["plain", "batch", "pbe", "adjoint", "adjoint_batch",
                                           "adjoint_general"],
                    default="plain")
    ap.add_argument("--ckpt-every", type=int, default=0,
                    help="adjoint checkpoint interval (0: the largest that fits in half of HBM)")
    ap.add_argument("--arith", choices=["fast", "exact"], default="fast")
    ap.add_argument("--pbe-arith", choices=["auto", "fast", "exact"], default="auto",
                    help="arithmetic form of the timed 2D workload (auto: exact)")
    ap.add_argument("--cells", type=int, default=1 << 24)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--variant", type=int, default=0,
                    help="1D FAST kernel: 0 register-blocked, 1 LDS tile (also the "
                         "adjoint_general leg's reverse pass)")
    ap.add_argument("--cells-per-lane", type=int, default=0,
                    help="cells per lane of the 1D stage kernel (0: chosen by the line length)")
    ap.add_argument("--per-line-steps", type=int, default=20)
    ap.add_argument("--legs", default="batch,adjoint,adjoint_batch,adjoint_general,pbe",
                    help="legs reported beside the headline ('none': the headline only)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain device copy of --cells doubles (20 times), the "
                         "known-size kernel that calibrates the FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path would "
              "shard by domain decomposition (lines of a batch: by line), which this library "
              "does not launch.", file=sys.stderr)
        return 2
    if args.steps is None:
        args.steps = LEG_STEPS[args.workload]

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

    runners = {"plain": run_plain, "batch": run_batch, "pbe": run_pbe, "adjoint": run_adjoint,
               "adjoint_batch": run_adjoint_batch, "adjoint_general": run_adjoint_general}
    value, ms_step, rl, config, data = runners[args.workload](args, torch, hh, W, dev)
    release(torch)

    cpu = None
    if not args.no_cpu_baseline and args.workload == "plain":
        log("cpu baseline ...")
        cpu = cpu_baseline(args.cpu_seconds)

    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": data, "config": config, "roofline": rl, "cpu_baseline": cpu,
    }
    legs = [] if args.legs in ("", "none") else [x.strip() for x in args.legs.split(",")]
    for leg in legs:
        if leg == args.workload:
            continue
        if leg not in runners:
            raise SystemExit(f"unknown leg {leg!r}")
        sub = argparse.Namespace(**vars(args))
        sub.workload = leg
        sub.steps = LEG_STEPS[leg]
        sub.ckpt_every = 0
        log(f"leg {leg}: {sub.steps} steps")
        v, ms, lrl, lconf, ldata = runners[leg](sub, torch, hh, W, dev)
        release(torch)
        out[leg] = {"value": v, "unit": "cell-updates/s", "steps": sub.steps,
                    "ms_per_step": ms, "config": lconf, "roofline": lrl, "data": ldata}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
