#!/usr/bin/env python3
"""Benchmark of the SPHENIX neighbour pass on one MI355X.

Workload (default): the GPU-scale drop-in case -- a periodic unit cube of
128^3 = 2,097,152 gas particles on a jittered lattice, displaced in x so that
rho ~ 1 + 0.5 sin(2 pi x), Gaussian velocities (sigma 0.1), u ~ U[0.5, 1.5],
all particles active, 32^3 top-level cells, Wendland C2, eta = 1.2348.

A step is one full pass over that box: density -> ghost (smoothing-length
iteration with subset re-runs) -> gradient -> extra ghost -> force -> end
force, starting each time from the same initial state (restored device to
device, inside the step).  The metric is SPH pair interactions/s counted as
SWIFT's counters count them: directed density interactions (j != i,
r < H_i) plus directed force interactions (r < max(H_i, H_j)) of the
converged pass; ghost re-runs are overhead.  value = batched interactions/s.

Beside it the line reports the plain per-task drop-in (one launch per
self/pair task, per-cell ghosts), p50/p99 step latency, the roofline of the
dominant kernel from HIP events on the engine stream, and the CPU baseline
(the oracle's FP32 restatement of the pass, threaded over the host cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic FP32 operations per directed interaction, counted from the
# reference's expressions (runner_iact_nonsym_* in
# src/hydro/SPHENIX/hydro_iact.h plus the pair distance and mask of the
# neighbour loop), an FMA-able multiply-add counted as 2, sqrt / divide /
# min / max / compare as 1, per-particle invariants (1/h, h^-4) excluded.
FLOP_DENSITY = 73
FLOP_GRADIENT = 61
FLOP_FORCE = 161
PEAK_FP32_TFLOPS = 157.3      # MI355X FP32 vector peak (packed issue)
PEAK_HBM_GBS = 8000.0


def _pct(a, q):
    return float(np.percentile(np.asarray(a), q))


def _read_traffic(path, kernel_key):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, copy-kernel scaled)."""
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        line = line.strip()
        if line.startswith(kernel_key + ":"):
            k, v = line[len(kernel_key) + 1:].split("=")
            vals[k.strip()] = float(v)
    if "hbm_bytes_per_launch" in vals:
        return vals["hbm_bytes_per_launch"]
    return None


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128, help="particles per axis")
    ap.add_argument("--ncell", type=int, default=32, help="top-level cells per axis")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads for the CPU baseline (0: min(16, cores))")
    ap.add_argument("--cpu-n", type=int, default=0,
                    help="particles per axis of the CPU sample (0: same as --n)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not "
              "supported (see DESIGN.md for how the pass would shard).",
              file=sys.stderr)
        return 2

    from swift_amd import Engine, HydroProps, ics

    st = ics.modulated_box(args.n, seed=0, amp=0.5)
    props = HydroProps(ncell=(args.ncell,) * 3)
    eng = Engine(props)
    eng.set_particles(st)
    eng.save_state()

    def step():
        eng.restore_state()
        return eng.pass_batched()

    for _ in range(args.warmup):
        step()
    print(f"bench: warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)
    eng.set_timing(True)
    eng.sync()
    times = []
    stats = None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        stats = step()          # returns after the pass has completed
        times.append(time.perf_counter() - ts)
    eng.sync()
    total = time.perf_counter() - t0
    ms, nl = eng.timing()
    print(f"bench: {args.steps} timed steps in {total:.3f} s", file=sys.stderr,
          flush=True)
    inter = stats["n_density"] + stats["n_force"]
    value = args.steps * inter / total

    # Roofline of the dominant neighbour-loop kernel (HIP events on the
    # engine stream over the timed region).
    names = ["density_loop", "gradient_loop", "force_loop"]
    dom = int(np.argmax(ms))
    kernels = {}
    for k in range(3):
        avg = ms[k] / max(nl[k], 1)
        entry = {"avg_ms": avg, "launches_per_step": nl[k] / args.steps,
                 "share_of_step": ms[k] / (1e3 * total)}
        if k == 0:
            # every density launch of the pass: the first round and the
            # ghost's re-runs, charged with the interactions they evaluated
            entry["interactions_evaluated"] = stats["n_density_all"]
            entry["tflops"] = (stats["n_density_all"] * FLOP_DENSITY * args.steps
                               / (ms[0] * 1e-3) / 1e12)
        if k == 1:
            entry["tflops"] = stats["n_gradient"] * FLOP_GRADIENT / (avg * 1e-3) / 1e12
        if k == 2:
            entry["tflops"] = stats["n_force"] * FLOP_FORCE / (avg * 1e-3) / 1e12
        kernels[names[k]] = entry
    if dom == 0:
        flops = kernels["density_loop"]["tflops"] * 1e12
        avg_ms = kernels["density_loop"]["avg_ms"]
    elif dom == 1:
        flops = kernels["gradient_loop"]["tflops"] * 1e12
        avg_ms = kernels["gradient_loop"]["avg_ms"]
    else:
        flops = kernels["force_loop"]["tflops"] * 1e12
        avg_ms = kernels["force_loop"]["avg_ms"]
    achieved = flops / 1e12
    traffic = _read_traffic(os.path.join(ROOT, "profiles", "pmc_traffic.txt"),
                            f"{names[dom]}_{args.n}")
    roofline = {"bound": "valu", "kernel": names[dom], "achieved": achieved,
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                "avg_launch_ms": avg_ms,
                "flop_per_interaction": [FLOP_DENSITY, FLOP_GRADIENT,
                                         FLOP_FORCE][dom]}

    out = {
        "metric": "SPH pair interactions/s (density+force) batched vs plain on GPU; p50/p99 step ms",
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "p50_step_ms": 1e3 * _pct(times, 50),
        "p99_step_ms": 1e3 * _pct(times, 99),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 positions)",
        "data": "synthetic",
        "config": {"workload": f"modulated periodic box {args.n}^3 "
                               f"({args.n ** 3} particles), rho ~ 1+0.5 sin(2 pi x), "
                               f"{args.ncell}^3 top-level cells, all active, "
                               "density+ghost+gradient+force",
                   "particles": args.n ** 3, "top_cells": args.ncell ** 3,
                   "parallelism": "single GPU"},
        "interactions_per_step": {"density": stats["n_density"],
                                  "force": stats["n_force"],
                                  "gradient": stats["n_gradient"]},
        "ghost_rounds": stats["ghost_rounds"],
        "launches_per_step": stats["launches"],
        "kernels": kernels,
        "roofline": roofline,
    }

    if not args.no_plain and args.plain_steps > 0:
        ptimes = []
        pstats = None
        for _ in range(args.plain_steps):
            eng.restore_state()
            eng.sync()
            ts = time.perf_counter()
            pstats = eng.pass_plain()
            eng.sync()
            ptimes.append(time.perf_counter() - ts)
        pinter = pstats["n_density"] + pstats["n_force"]
        pval = pinter / float(np.mean(ptimes))
        out["plain"] = {"value": pval, "unit": "interactions/s",
                        "ms_per_step": 1e3 * float(np.mean(ptimes)),
                        "launches_per_step": pstats["launches"],
                        "steps": args.plain_steps,
                        "same_counts": pinter == inter}
        out["batched_over_plain"] = value / pval
        print(f"bench: plain pass {np.mean(ptimes):.2f} s", file=sys.stderr,
              flush=True)
    eng.close()

    if not args.no_cpu:
        from oracle import oracle as orc
        ncpu = os.cpu_count() or 1
        threads = args.cpu_threads or min(16, ncpu)
        cn = args.cpu_n or args.n
        cst = ics.modulated_box(cn, seed=0, amp=0.5)
        cprops = HydroProps(ncell=(max(3, args.ncell * cn // args.n),) * 3)
        ts = time.perf_counter()
        o = orc.pass_fp32(cprops, cst, nthreads=threads)
        dt = time.perf_counter() - ts
        cinter = int(o["n_density"].astype(np.int64).sum()) + int(
            o["n_force"].astype(np.int64).sum())
        out["cpu_baseline"] = {
            "value": cinter / dt, "unit": "interactions/s", "cores": threads,
            "kind": "port",
            "sample": f"one full pass (density+ghost+gradient+force) over the "
                      f"{cn}^3 modulated box, {dt:.1f} s wall on {threads} "
                      "host threads (oracle FP32 restatement, OpenMP)"}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
