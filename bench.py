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


def _read_profile_value(path, kernel_key, name):
    """A `kernel_key: name = value` line of a committed rocprofv3 summary."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        line = line.strip()
        if line.startswith(kernel_key + ":"):
            k, v = line[len(kernel_key) + 1:].split("=")
            if k.strip() == name:
                return float(v)
    return None


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


def serving_bench(duration_s: float, rate: float, seed: int = 0,
                  max_inflight: int = 8, pool_size: int = 24) -> dict:
    """Poisson stream of independent boxes through the request server:
    N log-uniform in [16^3, 96^3] (rounded to a cube), each a modulated box
    with its own seed and amplitude U[0, 0.8], one density+force step per
    request.  Boxes come from a pool generated beforehand (generating one
    per request would time numpy, not the server)."""
    import threading
    from swift_amd import HydroProps, Server, ics
    rng = np.random.default_rng(seed)
    pool = []
    for k in range(pool_size):
        N = float(np.exp(rng.uniform(np.log(16 ** 3), np.log(96 ** 3))))
        n = int(round(N ** (1.0 / 3.0)))
        amp = float(rng.uniform(0.0, 0.8))
        st = ics.modulated_box(n, seed=1000 + k, amp=amp)
        # H <= cell width at the lowest density (rho = 1 - amp)
        props = HydroProps(ncell=(max(3, n // 5),) * 3)
        pool.append((props, st))
    print(f"bench: serving pool of {pool_size} boxes ready", file=sys.stderr,
          flush=True)
    infos = []
    ids = []
    lock = threading.Lock()
    fields = ["h", "rho", "pressure", "soundspeed", "v_sig", "alpha_visc",
              "alpha_diff", "a_hydro", "u_dt", "h_dt", "n_density", "n_force",
              "min_ngb_time_bin"]
    with Server(max_inflight=max_inflight, fields=fields) as srv:
        done = threading.Event()

        def collect():
            k = 0
            while True:
                with lock:
                    if k < len(ids):
                        rid = ids[k]
                    elif done.is_set():
                        return
                    else:
                        rid = None
                if rid is None:
                    time.sleep(0.0005)
                    continue
                infos.append(srv.wait(rid))
                srv.release(rid)
                k += 1

        th = threading.Thread(target=collect)
        th.start()
        t0 = time.perf_counter()
        t_next = t0
        while True:
            t_next += rng.exponential(1.0 / rate)
            if t_next - t0 > duration_s:
                break
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            props, st = pool[int(rng.integers(pool_size))]
            rid = srv.submit(props, st)
            with lock:
                ids.append(rid)
        done.set()
        th.join()
        wall = time.perf_counter() - t0
    lat = [i["t_done_ms"] - i["t_submit_ms"] for i in infos]
    queue = [i["t_admit_ms"] - i["t_submit_ms"] for i in infos]
    stages = {"upload_ms": [i["t_uploaded_ms"] - i["t_start_ms"] for i in infos],
              "pass_ms": [i["t_passed_ms"] - i["t_uploaded_ms"] for i in infos],
              "fetch_ms": [i["t_done_ms"] - i["t_passed_ms"] for i in infos]}
    inter = sum(i["n_density"] + i["n_force"] for i in infos)
    return {"workload": f"Poisson stream, mean {rate:g} req/s for {duration_s:g} s, "
                        "N log-uniform in [16^3, 96^3], modulated boxes with "
                        "amplitude U[0, 0.8], one density+force step each",
            "requests": len(infos), "max_inflight": max_inflight,
            "p50_latency_ms": _pct(lat, 50), "p99_latency_ms": _pct(lat, 99),
            "p50_queue_ms": _pct(queue, 50), "p99_queue_ms": _pct(queue, 99),
            "mean_stage_ms": {k: float(np.mean(v)) for k, v in stages.items()},
            "requests_per_s": len(infos) / wall,
            "interactions_per_s": inter / wall, "wall_s": wall}


def clustered_bench(n_side: int, ncell: int, steps: int, plain: bool,
                    seed: int = 0) -> dict:
    """The clustered, uneven workload: n_side^3 particles, 70 % in 1000
    Plummer spheres (scale radii log-uniform in [0.002, 0.02]) and 30 %
    uniform, u log-uniform in [0.1, 100], h from a converged all-active
    pre-pass, then 10 % of the particles active (random time-bin mask)."""
    from swift_amd import Engine, HydroProps, ics
    N = n_side ** 3
    t0 = time.perf_counter()
    st = ics.clustered_box(N, seed=seed, n_spheres=1000, a_min=0.002,
                           a_max=0.02)
    props = HydroProps(ncell=(ncell,) * 3)
    eng = Engine(props)
    eng.set_particles(st)
    pre = eng.pass_batched()
    res = eng.results(["h", "rho", "pressure", "soundspeed", "f", "balsara",
                       "alpha_visc", "alpha_diff"])
    for k, v in res.items():
        st[k] = v
    ics.set_active_fraction(st, 0.1, seed=seed + 1)
    eng.set_particles(st)
    eng.save_state()
    print(f"bench: clustered box ready ({time.perf_counter() - t0:.1f} s, "
          f"pre-pass {pre['ghost_rounds']} rounds)", file=sys.stderr,
          flush=True)
    tuning = eng.autotune(repeats=2)
    times = []
    for k in range(steps + 1):
        eng.restore_state()
        eng.sync()
        ts = time.perf_counter()
        s = eng.pass_graph()
        times.append(time.perf_counter() - ts)
    times = times[1:]  # the first replay captures the graphs
    inter = s["n_density"] + s["n_force"]
    h = res["h"]
    out = {"workload": f"clustered box {n_side}^3 ({N} particles), 1000 Plummer "
                       "spheres + 30 % uniform, h from a converged all-active "
                       f"pass, 10 % active, {ncell}^3 top-level cells",
           "particles": N, "active": int((st["time_bin"] <= 56).sum()),
           "h_max_over_min": float(h.max() / h.min()),
           "value": inter / float(np.mean(times)), "unit": "interactions/s",
           "ms_per_step": 1e3 * float(np.mean(times)),
           "p50_step_ms": 1e3 * _pct(times, 50),
           "p99_step_ms": 1e3 * _pct(times, 99),
           "interactions_per_step": {"density": s["n_density"],
                                     "force": s["n_force"]},
           "ghost_rounds": s["ghost_rounds"], "tuning": tuning}
    if plain:
        eng.restore_state()
        eng.sync()
        ts = time.perf_counter()
        ps = eng.pass_plain()
        eng.sync()
        pdt = time.perf_counter() - ts
        pinter = ps["n_density"] + ps["n_force"]
        out["plain"] = {"value": pinter / pdt, "unit": "interactions/s",
                        "ms_per_step": 1e3 * pdt,
                        "launches_per_step": ps["launches"]}
        out["batched_over_plain"] = out["value"] / (pinter / pdt)
    eng.close()
    return out


def _hbm_used_bytes() -> int:
    """Device memory in use (hipMemGetInfo through the HIP runtime)."""
    import ctypes as C
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        return -1
    free, total = C.c_size_t(), C.c_size_t()
    if hip.hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
        return -1
    return int(total.value - free.value)


def evolve_bench(n_side: int, ncell: int, nsteps: int, seed: int = 0,
                 chunk: int = 0) -> dict:
    """Multi-step run of the clustered construction: forces at t = 0 for all
    particles (the initial h estimate converged by the ghost), then `nsteps`
    drift-kick steps on a fixed 4-level time-bin hierarchy (bins 1..4 hold
    1.5 / 4.5 / 19 / 75 % of the particles, so 1.5, 6, 25 or 100 % are
    active), state resident in HBM, re-binned on the device when particles
    leave the stencil."""
    from swift_amd import Engine, HydroProps, ics
    from swift_amd.engine import integration
    N = n_side ** 3
    t0 = time.perf_counter()
    st = ics.clustered_box(N, seed=seed, n_spheres=1000, a_min=0.002,
                           a_max=0.02)
    ics.set_bin_hierarchy(st, {1: 0.015, 2: 0.045, 3: 0.19, 4: 0.75},
                          seed=seed + 3)
    t_ics = time.perf_counter() - t0
    print(f"bench: evolve box built ({t_ics:.1f} s)", file=sys.stderr, flush=True)
    # time line: the longest step (bin 4, 32 units) is 3.2e-8
    props = HydroProps(ncell=(ncell,) * 3, time_base=1e-9)
    eng = Engine(props)
    ts = time.perf_counter()
    eng.set_particles(st)
    eng.sync()
    t_up = time.perf_counter() - ts
    del st
    eng.set_integration(integration(fixed_bins=True))
    if chunk:
        for loop in ("density", "gradient", "force"):
            eng.set_chunk(loop, chunk)
    ts = time.perf_counter()
    first = eng.step()
    t_first = time.perf_counter() - ts
    print(f"bench: evolve step 0 (all active, {first['pass']['ghost_rounds']} "
          f"ghost rounds) {t_first:.2f} s", file=sys.stderr, flush=True)
    steps = []
    for k in range(nsteps):
        ts = time.perf_counter()
        info = eng.step()          # returns after the step has completed
        dt = time.perf_counter() - ts
        p = info["pass"]
        steps.append({"ti": info["ti_current"], "active": info["n_active"],
                      "ms": 1e3 * dt, "rebuilt": info["rebuilt"],
                      "sparse_tiles": info["sparse"],
                      "ghost_rounds": p["ghost_rounds"],
                      "interactions": p["n_density"] + p["n_force"]})
        print(f"bench: evolve step {k + 1}: {info['n_active']} active, "
              f"{1e3 * dt:.1f} ms", file=sys.stderr, flush=True)
    hbm = _hbm_used_bytes()
    eng.close()
    ms = [s["ms"] for s in steps]
    inter = sum(s["interactions"] for s in steps)
    wall = sum(ms) * 1e-3
    by_frac = {}
    for s in steps:
        f = round(100.0 * s["active"] / N, 1)
        by_frac.setdefault(f, []).append(s["ms"])
    return {
        "workload": f"clustered box {n_side}^3 ({N} particles), 1000 Plummer "
                    "spheres + 30 % uniform, u log-uniform [0.1, 100], "
                    f"{ncell}^3 top-level cells, {nsteps} drift-kick steps on "
                    "a fixed 4-level time-bin hierarchy (active 100/25/6/1.5 %),"
                    " state resident in HBM",
        "particles": N, "steps": nsteps,
        "value": inter / wall, "unit": "interactions/s",
        "ms_per_step": 1e3 * wall / nsteps,
        "p50_step_ms": _pct([m * 1e-3 for m in ms], 50) * 1e3,
        "p99_step_ms": _pct([m * 1e-3 for m in ms], 99) * 1e3,
        "ms_by_active_percent": {str(k): float(np.mean(v))
                                 for k, v in sorted(by_frac.items())},
        "rebuilds": sum(s["rebuilt"] for s in steps),
        "initial_step_s": t_first, "upload_s": t_up, "ics_s": t_ics,
        "hbm_used_gb": hbm / 1e9 if hbm >= 0 else None,
        "per_step": steps,
    }


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
    ap.add_argument("--cpu-n1", type=int, default=0,
                    help="particles per axis of the single-thread CPU sample "
                         "(0: half of the threaded sample's)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default schedule and tile shapes")
    ap.add_argument("--graph", type=int, default=1,
                    help="1: replay each step from captured HIP graphs")
    ap.add_argument("--serving-s", type=float, default=5.0,
                    help="seconds of the serving stream (0: skip)")
    ap.add_argument("--serving-rate", type=float, default=200.0)
    ap.add_argument("--clustered-n", type=int, default=256,
                    help="particles per axis of the clustered workload (0: skip)")
    ap.add_argument("--clustered-ncell", type=int, default=48)
    ap.add_argument("--evolve-n", type=int, default=0,
                    help="particles per axis of the multi-step clustered run "
                         "(0: skip; 512 is the full-size case)")
    ap.add_argument("--evolve-ncell", type=int, default=96)
    ap.add_argument("--evolve-steps", type=int, default=20)
    ap.add_argument("--evolve-chunk", type=int, default=0,
                    help="chunk capacity of all loops in the multi-step run "
                         "(0: the engine's choice)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not "
              "supported: the pass shards by space over top-level cells with "
              "one-cell halos, which is not built here.",
              file=sys.stderr)
        return 2

    from swift_amd import Engine, HydroProps, ics

    st = ics.modulated_box(args.n, seed=0, amp=0.5)
    props = HydroProps(ncell=(args.ncell,) * 3)
    eng = Engine(props)
    eng.set_particles(st)
    eng.save_state()
    tuning = None
    if not args.no_tune:
        tuning = eng.autotune(repeats=2)
        print(f"bench: autotuned {tuning}", file=sys.stderr, flush=True)
    run = eng.pass_graph if args.graph else eng.pass_batched

    def step():
        eng.restore_state()
        return run()

    for _ in range(args.warmup):
        step()
    print(f"bench: warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)
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
    print(f"bench: {args.steps} timed steps in {total:.3f} s", file=sys.stderr,
          flush=True)
    inter = stats["n_density"] + stats["n_force"]
    value = args.steps * inter / total

    # Per-loop kernel times: HIP events recorded on the engine stream around
    # every neighbour-loop kernel launch (the tile build excluded) of the
    # same number of eager passes with the same kernels (a captured graph
    # carries no per-launch events).
    eng.set_timing(True)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        eng.restore_state()
        eager_stats = eng.pass_batched()
    eng.sync()
    eager_total = time.perf_counter() - t1
    ms, nl = eng.timing()
    eng.set_timing(False)
    assert eager_stats["n_density"] == stats["n_density"]

    # Roofline of the dominant neighbour-loop kernel.
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
    valu_busy = _read_profile_value(
        os.path.join(ROOT, "profiles", "pmc_valu_128cube.md"),
        f"{names[dom]}_{args.n}", "valu_busy")
    roofline = {"bound": "valu", "kernel": names[dom], "achieved": achieved,
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                "hbm_gbs": (traffic / (avg_ms * 1e-3) / 1e9) if traffic else None,
                "valu_busy": valu_busy,
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
        "pass": "graph replay" if args.graph else "eager batched",
        "eager_ms_per_step": 1e3 * eager_total / args.steps,
        "tuning": tuning,
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
                        # the plain task order sums in another order, so h
                        # and borderline neighbour counts may differ slightly
                        "same_counts": pinter == inter,
                        "count_rel_diff": abs(pinter - inter) / max(inter, 1)}
        out["batched_over_plain"] = value / pval
        print(f"bench: plain pass {np.mean(ptimes):.2f} s", file=sys.stderr,
              flush=True)
    eng.close()

    if args.clustered_n > 0:
        out["clustered"] = clustered_bench(args.clustered_n, args.clustered_ncell,
                                           args.steps, not args.no_plain)
        print(f"bench: clustered {out['clustered']['ms_per_step']:.1f} ms/step",
              file=sys.stderr, flush=True)

    if args.evolve_n > 0:
        out["evolve"] = evolve_bench(args.evolve_n, args.evolve_ncell,
                                     args.evolve_steps, chunk=args.evolve_chunk)
        print(f"bench: evolve {out['evolve']['ms_per_step']:.1f} ms/step",
              file=sys.stderr, flush=True)

    if args.serving_s > 0:
        out["serving"] = serving_bench(args.serving_s, args.serving_rate)
        print(f"bench: serving {out['serving']['requests']} requests",
              file=sys.stderr, flush=True)

    if not args.no_cpu:
        from oracle import oracle as orc
        ncpu = os.cpu_count() or 1
        threads = args.cpu_threads or min(16, ncpu)
        cn = args.cpu_n or args.n

        def cpu_pass(n_side, nthreads):
            cs = ics.modulated_box(n_side, seed=0, amp=0.5)
            cp = HydroProps(ncell=(max(3, args.ncell * n_side // args.n),) * 3)
            ts = time.perf_counter()
            o = orc.pass_fp32(cp, cs, nthreads=nthreads, port=True)
            dt = time.perf_counter() - ts
            ci = int(o["n_density"].astype(np.int64).sum()) + int(
                o["n_force"].astype(np.int64).sum())
            return ci / dt, dt

        cval, dt = cpu_pass(cn, threads)
        cn1 = args.cpu_n1 or max(8, cn // 2)
        cval1, dt1 = cpu_pass(cn1, 1)
        out["cpu_baseline"] = {
            "value": cval, "unit": "interactions/s", "cores": threads,
            "kind": "port",
            "sample": f"one full pass (density+ghost+gradient+force) over the "
                      f"{cn}^3 modulated box, {dt:.1f} s wall on {threads} "
                      "host threads (the oracle's FP32 restatement with sorted "
                      "per-axis candidate lists, -O3 -ffast-math, OpenMP over "
                      "cells)",
            "single_thread": {"value": cval1, "unit": "interactions/s",
                              "sample": f"{cn1}^3 modulated box, {dt1:.1f} s "
                                        "wall on 1 thread"}}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
