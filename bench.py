"""Benchmark of the MIND simulation step on one MI355X.

Workload: 5000 seed-generated branched cells (compartment counts lognormal,
median 400, clipped to 50..4000; about 2.7 M compartments and 5.4 M cable
nodes), hh soma, pas dendrites, 20 Exp2Syn per cell, in-degree 50 recurrent
NetCons and a 5 Hz Poisson NetStim per cell, dt 0.025 ms, FP64.

A step is one micro step (dt) of the whole network: detection, delivery,
nrn_cur, cable solve, nrn_state, recording. `value` is compartment-steps/s of
the batched kernel with the model resident in HBM (build, autotune and
initialize are outside the timed region); the plain one-cell-per-thread kernel
is timed on the same model beside it.

Also reported: the roofline of the dominant kernel (algorithmic bytes of its
launch over its average duration measured with HIP events on the library's
stream) and the CPU oracle timed on the host on a bounded sample of the same
workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mind.jl_amd"))

HBM_PEAK_GBS = 8000.0
# Compulsory bytes per step of the micro step's state, touched once:
# node v r/w 16 + cap 8 + axial 8 + axial diagonal 8 + area 8 + parent 4;
# pas g, e; hh 4 parameters + m, h, n r/w + ena, ek; Exp2Syn tau1, tau2, e,
# factor + A, B r/w + node; IClamp; detector state.
BYTES = {"nodes": 52, "pas": 16, "hh": 96, "exp2syn": 68, "iclamp": 28, "detectors": 2}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cells", type=int, default=5000)
    ap.add_argument("--plain-steps", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--macro-ms", type=float, default=2000.0)
    ap.add_argument("--no-macro", action="store_true")
    ap.add_argument("--plan", default="", metavar="BUDGET,THREADS,LAYOUT",
                    help="batched plan instead of the autotuner's (e.g. to profile a given plan)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the plain streaming copy of BYTES (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import ctypes as C

        from mind_amd import _abi

        g = C.c_double()
        _abi.check(_abi.lib().mind_copy_bandwidth(0, args.copy_calibration, 20, C.byref(g)))
        print(json.dumps({"copy_bytes": args.copy_calibration, "repeats": 20, "copy_gbps": g.value}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the sharded path is "
              "design-only", file=sys.stderr)
        sys.exit(2)

    from mind_amd import build, configs, initialize, run, destroy
    from mind_amd.model import (batch_plan, batch_work, kernel_time, set_batch_plan, set_profiling,
                                synchronize, window_steps)

    dt = 0.025
    t0 = time.time()
    model, info = configs.tree_network(ncells=args.cells)
    log(f"generated model in {time.time() - t0:.1f}s; compartments {int(info['sizes'].sum())}")

    def timed(method, steps, warmup, profile):
        t1 = time.time()
        build(model, dt_micro=dt, method=method)
        if method == "batched" and args.plan:
            set_batch_plan(model, *(int(x) for x in args.plan.split(",")))
        initialize(model)
        log(f"{method}: build+autotune+initialize {time.time() - t1:.1f}s")
        if warmup:
            run(model, warmup * dt)
        set_profiling(model, profile)
        synchronize(model)
        t2 = time.perf_counter()
        res = run(model, (warmup + steps) * dt)  # returns after its stream has drained
        synchronize(model)
        wall = time.perf_counter() - t2
        kt = kernel_time(model) if profile else None
        extra = {}
        if method == "batched":
            extra["plan"] = batch_plan(model)
            extra["work"] = batch_work(model)
        extra["window"] = window_steps(model)
        extra["spikes"] = int(len(res.spike_times))
        return wall, kt, extra

    import ctypes as C

    from mind_amd import _abi

    wall, kt, extra = timed("batched", args.steps, args.warmup, True)
    n, s = C.c_int64(), C.c_int64()
    _abi.lib().mind_node_count(model.state, C.byref(n), C.byref(s))
    nodes, comps = n.value, s.value
    value = comps * args.steps / wall
    launches, ksteps, kms = kt
    work = extra["work"]
    bytes_per_step = sum(BYTES[k] * work[k] for k in BYTES)
    bytes_per_launch = bytes_per_step * (ksteps / launches)
    avg_launch_s = kms / launches / 1e3
    achieved = bytes_per_launch / avg_launch_s / 1e9
    kernel = {0: "k_window", 4: "k_window_lean_hbm", 5: "k_window_lean"}[extra["plan"]["layout"]]
    traffic, traffic_note = measured_traffic(kernel, extra["plan"], extra["window"], ksteps / launches,
                                             bytes_per_launch)
    destroy(model)

    plain = None
    if not args.no_plain:
        pw, _, pextra = timed("plain", args.plain_steps, 0, False)
        plain = {"value": comps * args.plain_steps / pw, "steps": args.plain_steps,
                 "ms_per_step": 1e3 * pw / args.plain_steps}
        destroy(model)

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        sim = oracle.OracleSim(model, dt=dt).initialize()
        r = sim.run(args.cpu_steps * dt, record_spikes=False)
        cpu = {"value": comps * args.cpu_steps / r.run_seconds, "unit": "compartment-steps/s",
               "cores": 1, "kind": "port",
               "sample": f"the same {args.cells}-cell model, first {args.cpu_steps} steps "
                         f"after initialize, oracle (C++ FP64, -O3 -ffp-contract=off), 1 thread"}

    macro = None if args.no_macro else macro_leg(args.macro_ms)

    line = {
        "metric": "compartment-steps/s (batched vs plain); macro sim-s/wall-s surrogate vs plain",
        "value": value,
        "unit": "compartment-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.cells} seed-generated branched cells (lognormal median 400 "
                        "compartments), hh/pas/Exp2Syn, in-degree 50, 5 Hz NetStim; batched "
                        "window kernel",
            "cells": args.cells, "compartments": comps, "nodes": nodes, "dt_ms": dt,
            "window_steps": extra["window"], "plan": extra["plan"],
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_source": traffic_note,
            "kernel": kernel, "bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": kms / launches, "launches": launches,
        },
        "plain": plain,
        "batched_over_plain": (value / plain["value"]) if plain else None,
        "cpu_baseline": cpu,
        "sim_seconds_per_wall_second": (args.steps * dt / 1e3) / wall,
        "spikes_in_timed_run": extra["spikes"],
        "macro": macro,
    }
    print(json.dumps(line), flush=True)


def measured_traffic(kernel, plan, window, steps_per_launch, algorithmic_per_launch):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summaries (profiles/*.md, 'bench-traffic:' lines),
    scaled from a full window to this run's average launch. A summary of the same
    plan is used as measured; one of the same kernel with another plan (the
    autotuner's pick can change between runs) through its ratio of measured to
    algorithmic bytes, and the returned note says so. (None, None) without one."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None, None
    near = None
    for f in sorted(os.listdir(d)):
        if not f.endswith(".md"):
            continue
        for line in open(os.path.join(d, f)):
            if not line.startswith("bench-traffic: "):
                continue
            rec = json.loads(line[len("bench-traffic: "):])
            if rec["kernel"] != kernel or rec["window_steps"] != window:
                continue
            if rec["plan"] == plan:
                return rec["hbm_bytes_per_window"] * steps_per_launch / window, f"profiles/{f}"
            if near is None and rec.get("algorithmic_bytes_per_window"):
                near = (rec, f)
    if near is None:
        return None, None
    rec, f = near
    ratio = rec["hbm_bytes_per_window"] / rec["algorithmic_bytes_per_window"]
    return ratio * algorithmic_per_launch, f"profiles/{f}, plan {rec['plan']}, scaled by measured/algorithmic bytes"


def macro_leg(stop):
    """16 Wilson-Cowan ROIs x 256 ball-and-stick cells (evaluation seed 1): the
    detailed micro-macro run (batched micro kernel) against the surrogate run of
    the same ROI graph, in simulated s per wall s of the run loop."""
    import numpy as np

    from mind_amd import configs, destroy
    from mind_amd import macro as M
    from mind_amd import surrogate as S

    mm, _ = configs.micro_macro(seed=1)
    t0 = time.time()
    M.build(mm, seed=1, method="batched")
    M.initialize(mm)
    log(f"micro-macro: build+autotune+initialize {time.time() - t0:.1f}s")
    plain = M.run(mm, stop, record_spikes=False)
    log(f"micro-macro: {stop:.0f} ms in {plain.run_seconds:.2f}s")
    destroy(mm)
    mm.micro.state = None
    out = {"workload": "16 Wilson-Cowan ROIs (RK4, dt 0.1 ms, delayed all-to-all coupling) x 256 "
                       "200-compartment ball-and-stick cells, Poisson drive, 1 ms binned rate; seed 1",
           "stop_ms": stop, "plain_sim_s_per_wall_s": stop / 1e3 / plain.run_seconds,
           "plain_run_s": plain.run_seconds}
    if not os.path.exists(S.WEIGHTS):
        out["surrogate"] = None
        return out
    net = S.Network.load()
    cs = S.CoupledSurrogate(mm, net).initialize()
    sur = cs.run(stop)
    cs.close()
    rp = plain.rate[:, 10::10].mean(axis=1)
    rs = sur.result.rate[:, 10::10].mean(axis=1)
    err = np.abs(rs - rp) / rp
    verr = np.abs(sur.result.mean_v[:, 10::10].mean(axis=1) - plain.mean_v[:, 10::10].mean(axis=1))
    out.update({
        "surrogate_sim_s_per_wall_s": stop / 1e3 / sur.result.run_seconds,
        "surrogate_run_s": sur.result.run_seconds,
        "speedup": plain.run_seconds / sur.result.run_seconds,
        "rate_rel_err_max": float(err.max()), "rate_rel_err_mean": float(err.mean()),
        "mean_v_abs_err_max_mV": float(verr.max()),
        "ood_windows": sur.ood_windows,
    })
    log(f"surrogate: {stop:.0f} ms in {sur.result.run_seconds:.4f}s; rate error max {err.max():.4f}")
    return out


if __name__ == "__main__":
    main()
