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

The timed region runs on the library's normal path (full windows replayed as
a hipGraph, no per-launch timing) after an untimed warm-up of at least 80
steps, so that events of the 1 ms minimum delay are delivered inside it; the
events delivered there are reported.

Also reported: the roofline of the dominant kernel (algorithmic bytes of its
launch over its average duration, measured with HIP events on the library's
stream in a separate profiled pass after the timed region); the CPU oracle
timed on the host on a bounded sample of the same workload, on one thread and
on every usable core; the 1000-cell ball-and-stick network over its whole
1000 ms in both kernels; and the micro-macro coupling, detailed against the
learned surrogate.
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
    ap.add_argument("--no-config1", action="store_true")
    ap.add_argument("--eval-seeds", type=int, default=16,
                    help="closed-loop surrogate evaluation over seeds 1..N of the micro-macro workload")
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

    from mind_amd import configs, destroy

    dt = 0.025
    t0 = time.time()
    model, info = configs.tree_network(ncells=args.cells)
    log(f"generated model in {time.time() - t0:.1f}s; compartments {int(info['sizes'].sum())}")
    warm = max(args.warmup, MIN_WARMUP_STEPS)
    b = timed_leg(model, "batched", args.steps, warm, plan=args.plan, profile_windows=3)
    comps, nodes = b["compartments"], b["nodes"]
    value = comps * args.steps / b["wall"]
    launches, ksteps, kms = b["kernel_time"]
    work = b["work"]
    bytes_per_step = sum(BYTES[k] * work[k] for k in BYTES)
    bytes_per_launch = bytes_per_step * (ksteps / launches)
    avg_launch_s = kms / launches / 1e3
    achieved = bytes_per_launch / avg_launch_s / 1e9
    kernel = b["kernel"]
    traffic, traffic_note = measured_traffic(kernel, b["plan"], b["window"], ksteps / launches,
                                             bytes_per_launch)
    destroy(model)

    plain = None
    if not args.no_plain:
        pl = timed_leg(model, "plain", args.plain_steps, warm)
        plain = {"value": comps * args.plain_steps / pl["wall"], "steps": args.plain_steps,
                 "ms_per_step": 1e3 * pl["wall"] / args.plain_steps, "warmup": warm,
                 "events_delivered": pl["delivered"]}
        destroy(model)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_leg(model, comps, dt, skip=MIN_WARMUP_STEPS, steps_one=args.cpu_steps,
                      steps_all=4 * args.cpu_steps,
                      what=f"the same {args.cells}-cell model")

    config1 = None if args.no_config1 else config1_leg(dt, args.no_plain, args.no_cpu)
    macro = None if args.no_macro else macro_leg(args.macro_ms, args.eval_seeds)

    line = {
        "metric": "compartment-steps/s (batched vs plain); macro sim-s/wall-s surrogate vs plain",
        "value": value,
        "unit": "compartment-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * b["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.cells} seed-generated branched cells (lognormal median 400 "
                        "compartments), hh/pas/Exp2Syn, in-degree 50, 5 Hz NetStim; batched "
                        "window kernel replayed as a hipGraph",
            "cells": args.cells, "compartments": comps, "nodes": nodes, "dt_ms": dt,
            "window_steps": b["window"], "plan": b["plan"],
        },
        "timed_region": {
            "warmup_steps_run": warm,
            "warmup_note": (f"untimed warm-up extended from {args.warmup} to {warm} steps so that "
                            "events of the 1 ms minimum delay are delivered in the timed region")
                           if warm > args.warmup else "as requested",
            "start_ms": warm * dt, "end_ms": (warm + args.steps) * dt,
            "events_delivered": b["delivered"], "events_routed": b["routed"],
            "spikes": b["spikes"], "profiling": False,
            "graph_replay": "every full window of the run" if args.steps >= b["window"]
                            else f"no: {args.steps} steps are less than one {b['window']}-step window",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_source": traffic_note,
            "kernel": kernel, "bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": kms / launches, "launches": launches,
            "timing": f"HIP events around each {kernel} launch on the library's stream, in a "
                      f"separate profiled pass of {ksteps} steps after the timed region",
        },
        "plain": plain,
        "batched_over_plain": (value / plain["value"]) if plain else None,
        "cpu_baseline": cpu,
        "sim_seconds_per_wall_second": (args.steps * dt / 1e3) / b["wall"],
        "config1": config1,
        "macro": macro,
    }
    print(json.dumps(line), flush=True)


# At least two 1 ms windows before timing: by then NetStim and recurrent events
# (minimum delay 1 ms) are arriving every step.
MIN_WARMUP_STEPS = 80


def timed_leg(model, method, steps, warm, plan="", profile_windows=0):
    """Build and initialize (untimed), `warm` untimed steps, then `steps` timed on
    the library's normal path (hipGraph replay of full windows, profiling off);
    with profile_windows, that many more windows with HIP-event timing of the
    dominant launch."""
    import ctypes as C

    from mind_amd import _abi, build, initialize, run
    from mind_amd.model import (batch_kernel, batch_plan, batch_work, event_counts, kernel_time, set_batch_plan,
                                set_profiling, synchronize, window_steps)

    dt = 0.025
    t1 = time.time()
    build(model, dt_micro=dt, method=method)
    if method == "batched" and plan:
        set_batch_plan(model, *(int(x) for x in plan.split(",")))
    initialize(model)
    out = {"setup_s": time.time() - t1}
    log(f"{method}: build+autotune+initialize {out['setup_s']:.1f}s")
    if warm:
        run(model, warm * dt)
    set_profiling(model, False)
    r0, d0 = event_counts(model)
    synchronize(model)
    t2 = time.perf_counter()
    res = run(model, (warm + steps) * dt, record_spikes=True)  # returns after its stream has drained
    synchronize(model)
    out["wall"] = time.perf_counter() - t2
    r1, d1 = event_counts(model)
    out.update(routed=r1 - r0, delivered=d1 - d0, spikes=int(len(res.spike_times)), window=window_steps(model))
    n, s = C.c_int64(), C.c_int64()
    _abi.lib().mind_node_count(model.state, C.byref(n), C.byref(s))
    out.update(nodes=n.value, compartments=s.value)
    if method == "batched":
        out["plan"] = batch_plan(model)
        out["kernel"] = batch_kernel(model)
        out["work"] = batch_work(model)
    if profile_windows:
        set_profiling(model, True)
        run(model, (warm + steps + profile_windows * out["window"]) * dt, record_spikes=False)
        out["kernel_time"] = kernel_time(model)
        set_profiling(model, False)
    return out


def cpu_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count(), len(os.sched_getaffinity(0))


def cpu_leg(model, comps, dt, skip, steps_one, steps_all, what):
    """The oracle (C++ FP64 restatement, -O3 -ffp-contract=off) on the host: `skip`
    untimed steps so that deliveries are under way, then `steps_one` steps on one
    thread and `steps_all` more with OpenMP over cells on every usable core (the
    event queue stays serial). `value` is the all-core rate."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    threads = oracle.default_threads()
    sim = oracle.OracleSim(model, dt=dt, threads=threads).initialize()
    sim.run(skip * dt, record_spikes=False)
    sim.set_threads(1)
    r1 = sim.run((skip + steps_one) * dt, record_spikes=False)
    sim.set_threads(threads)
    rn = sim.run((skip + steps_one + steps_all) * dt, record_spikes=False)
    cpu_model, nproc, usable = cpu_info()
    one = comps * steps_one / r1.run_seconds
    return {"value": comps * steps_all / rn.run_seconds, "unit": "compartment-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{what}, steps {skip + steps_one}-{skip + steps_one + steps_all} after initialize "
                      f"(deliveries under way), oracle (C++ FP64, -O3 -ffp-contract=off), OpenMP over cells "
                      f"on {threads} threads, event queue serial",
            "one_thread": {"value": one, "steps": steps_one, "first_step": skip},
            "nproc": nproc, "usable_cpus": usable, "cpu_model": cpu_model}


def config1_leg(dt, no_plain, no_cpu):
    """The 1000-cell ball-and-stick network (in-degree 100, 10 Hz Poisson drive)
    over its whole 1000 ms in both kernels, and the oracle on a sample of it."""
    from mind_amd import configs, destroy

    model, info = configs.ball_stick_network()
    stop_steps = int(round(info["tstop"] / dt))
    out = {"workload": "1000 ball-and-stick cells (hh soma, 50-segment pas dendrite, Exp2Syn), in-degree "
                       "100, lognormal weights, delays U(1,5) ms, 10 Hz Poisson NetStim per cell; 1000 ms",
           "stop_ms": info["tstop"]}
    legs = ["batched"] + ([] if no_plain else ["plain"])
    for method in legs:
        r = timed_leg(model, method, stop_steps, 0)
        out[method] = {"value": r["compartments"] * stop_steps / r["wall"], "unit": "compartment-steps/s",
                       "run_s": r["wall"], "sim_s_per_wall_s": info["tstop"] / 1e3 / r["wall"],
                       "spikes": r["spikes"], "events_delivered": r["delivered"],
                       "plan": r.get("plan")}
        out["compartments"] = r["compartments"]
        log(f"config-1 {method}: {info['tstop']:.0f} ms in {r['wall']:.2f}s")
        destroy(model)
    if "plain" in out:
        out["batched_over_plain"] = out["batched"]["value"] / out["plain"]["value"]
    if not no_cpu:
        out["cpu_baseline"] = cpu_leg(model, out["compartments"], dt, skip=MIN_WARMUP_STEPS, steps_one=1000,
                                      steps_all=4000, what="the 1000-cell network")
    return out


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


def macro_leg(stop, eval_seeds):
    """16 Wilson-Cowan ROIs x 256 ball-and-stick cells. The detailed micro-macro run
    (batched micro kernel, one population-rate bin per window) against the
    surrogate run of the same ROI graph for evaluation seed 1, in simulated s per
    wall s of the run loop; then, closed loop on evaluation seeds 1..eval_seeds,
    the per-ROI time-averaged rate error of the surrogate and of its linear map
    alone (the method without the network)."""
    import numpy as np

    from mind_amd import configs, destroy
    from mind_amd import macro as M
    from mind_amd import surrogate as S

    def detailed(seed):
        mm, _ = configs.micro_macro(seed=seed)
        t0 = time.time()
        M.build(mm, seed=1, method="batched")
        M.initialize(mm)
        setup = time.time() - t0
        res = M.run(mm, stop, record_spikes=False)
        destroy(mm)
        mm.micro.state = None
        return mm, res, setup

    def errors(plain, sur):
        rp = plain.rate[:, 10::10].mean(axis=1)
        rs = sur.rate[:, 10::10].mean(axis=1)
        verr = np.abs(sur.mean_v[:, 10::10].mean(axis=1) - plain.mean_v[:, 10::10].mean(axis=1))
        return np.abs(rs - rp) / rp, verr

    mm, plain, setup = detailed(1)
    log(f"micro-macro: build+autotune+initialize {setup:.1f}s")
    log(f"micro-macro: {stop:.0f} ms in {plain.run_seconds:.2f}s")
    out = {"workload": "16 Wilson-Cowan ROIs (RK4, dt 0.1 ms, delayed all-to-all coupling) x 256 "
                       "200-compartment ball-and-stick cells, Poisson drive, 1 ms binned rate; seed 1",
           "stop_ms": stop, "plain_sim_s_per_wall_s": stop / 1e3 / plain.run_seconds,
           "plain_run_s": plain.run_seconds,
           "plain_windows": "one 1 ms rate bin (10 macro steps, 40 micro steps) per replayed window"}
    if not os.path.exists(S.WEIGHTS):
        out["surrogate"] = None
        return out
    net = S.Network.load()
    lin = S.linear_only(net)
    cs = S.CoupledSurrogate(mm, net).initialize()
    sur = cs.run(stop)
    cs.close()
    err, verr = errors(plain, sur.result)
    out.update({
        "surrogate_sim_s_per_wall_s": stop / 1e3 / sur.result.run_seconds,
        "surrogate_run_s": sur.result.run_seconds,
        "speedup": plain.run_seconds / sur.result.run_seconds,
        "rate_rel_err_max": float(err.max()), "rate_rel_err_mean": float(err.mean()),
        "mean_v_abs_err_max_mV": float(verr.max()),
        "ood_windows": sur.ood_windows,
    })
    log(f"surrogate: {stop:.0f} ms in {sur.result.run_seconds:.4f}s; rate error max {err.max():.4f}")
    net_err, lin_err, ood = [], [], 0
    for seed in range(1, eval_seeds + 1):
        if seed == 1:
            mms, ps = mm, plain
        else:
            mms, ps, _ = detailed(seed)
        for model_net, acc in ((net, net_err), (lin, lin_err)):
            cs = S.CoupledSurrogate(mms, model_net).initialize()
            r = cs.run(stop)
            cs.close()
            if model_net is net:
                ood += r.ood_windows
            acc.append(errors(ps, r.result)[0])
        log(f"seed {seed}: rate error network max {net_err[-1].max():.4f}, linear map {lin_err[-1].max():.4f}")
    ne, le = np.concatenate(net_err), np.concatenate(lin_err)
    out["closed_loop_eval"] = {
        "seeds": f"1-{eval_seeds}", "stop_ms": stop, "rois": int(ne.size),
        "network_rate_rel_err": {"max": float(ne.max()), "mean": float(ne.mean())},
        "linear_map_rate_rel_err": {"max": float(le.max()), "mean": float(le.mean())},
        "network_beats_linear_map_mean": bool(ne.mean() < le.mean()),
        "seeds_within_5pct": int(sum(float(e.max()) <= 0.05 for e in net_err)),
        "ood_windows": int(ood),
    }
    return out


if __name__ == "__main__":
    main()
