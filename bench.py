#!/usr/bin/env python3
"""Benchmark of the SPHENIX neighbour pass on one MI355X.

Headline workload: the clustered, uneven box -- 256^3 = 16,777,216 gas
particles in a periodic unit cube, 70 % drawn from 1000 Plummer spheres
(centres uniform, scale radii log-uniform in [0.002, 0.02], member counts
Poisson) and 30 % uniform, equal masses, u log-uniform in [0.1, 100]; h
converged by an all-active pass, then 10 % of the particles active by a
random time-bin mask; 48^3 top-level cells (1 to ~5000 particles per cell,
h spread over ~80x).  The batched pass runs the device work queue with
autotuned tile shapes and is replayed from captured HIP graphs.

A step is one full pass over that box: density -> ghost (smoothing-length
iteration with subset re-runs) -> gradient -> extra ghost -> force -> end
force, starting each time from the same state (restored device to device,
inside the step).  The metric is SPH pair interactions/s counted as SWIFT's
counters count them: directed density interactions (j != i, r < H_i) plus
directed force interactions (r < max(H_i, H_j)) of the converged pass;
ghost re-runs are overhead.  value = batched interactions/s.

Beside it the line reports the plain per-task drop-in (one launch per
self/pair task, per-cell ghosts), p50/p99 step latency, the roofline of the
dominant kernel from HIP events on the engine stream, the CPU baseline (the
oracle's FP32 restatement of the pass with sorted candidate lists, built
with -march=native on this host, threaded and pinned over the host cores,
median of 5 after a warm-up, on the same clustered state), and side lines
for the all-active modulated 128^3 box, the serving stream and (optional)
the 512^3 multi-step run.
"""
from __future__ import annotations

import os

# The CPU baseline's OpenMP threads are pinned one per core; libgomp reads
# these when it is first loaded, so they are set before anything loads it.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import argparse  # noqa: E402
import json  # noqa: E402
import subprocess  # noqa: E402
import sys  # noqa: E402
import time  # noqa: E402

import numpy as np  # noqa: E402

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
LOOPS = ["density_loop", "gradient_loop", "force_loop"]
IO_FIELDS = ["h", "rho", "pressure", "soundspeed", "f", "balsara",
             "alpha_visc", "alpha_diff"]


def _pct(a, q):
    return float(np.percentile(np.asarray(a), q))


def _profile_values(key):
    """`key: name = value` lines of the committed rocprofv3 summaries in
    profiles/ (the figures a summary reports for a kernel and workload)."""
    vals = {}
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return vals
    for fn in sorted(os.listdir(pdir)):
        if not fn.endswith((".md", ".txt")):
            continue
        for line in open(os.path.join(pdir, fn)):
            line = line.strip()
            if line.startswith(key + ":"):
                k, v = line[len(key) + 1:].split("=")
                vals[k.strip()] = float(v)
    return vals


def clustered_state(n_side: int, ncell: int, seed: int = 0):
    """The clustered box with h and the (io) state of a converged all-active
    pass, then 10 % of the particles active."""
    from swift_amd import Engine, HydroProps, ics
    N = n_side ** 3
    t0 = time.perf_counter()
    st = ics.clustered_box(N, seed=seed, n_spheres=1000, a_min=0.002,
                           a_max=0.02)
    props = HydroProps(ncell=(ncell,) * 3)
    with Engine(props) as eng:
        eng.set_particles(st)
        pre = eng.pass_batched()
        res = eng.results(IO_FIELDS)
    for k, v in res.items():
        st[k] = v
    ics.set_active_fraction(st, 0.1, seed=seed + 1)
    print(f"bench: clustered {n_side}^3 box ready ({time.perf_counter() - t0:.1f} s,"
          f" pre-pass {pre['ghost_rounds']} rounds)", file=sys.stderr, flush=True)
    workload = (f"clustered box {n_side}^3 ({N} particles): 70 % in 1000 Plummer "
                "spheres (scale radii log-uniform [0.002, 0.02]) + 30 % uniform, "
                "u log-uniform [0.1, 100], h from a converged all-active pass, "
                f"10 % active (random time-bin mask), {ncell}^3 top-level cells, "
                "density+ghost+gradient+force")
    return props, st, workload


def modulated_state(n_side: int, ncell: int):
    from swift_amd import HydroProps, ics
    st = ics.modulated_box(n_side, seed=0, amp=0.5)
    props = HydroProps(ncell=(ncell,) * 3)
    workload = (f"modulated periodic box {n_side}^3 ({n_side ** 3} particles), "
                f"rho ~ 1+0.5 sin(2 pi x), {ncell}^3 top-level cells, all active, "
                "density+ghost+gradient+force")
    return props, st, workload


def measure(props, st, workload: str, steps: int, warmup: int, tune: bool,
            graph: bool, plain: bool, traffic_key: str, shape=None) -> dict:
    """Times `steps` passes (after `warmup`), then the per-loop kernel times
    of as many eager passes with HIP events on the engine stream, then one
    plain per-task pass.  Returns the measurement and the roofline of the
    dominant loop."""
    from swift_amd import Engine
    eng = Engine(props)
    eng.set_particles(st)
    eng.save_state()
    tuning = None
    if shape:
        # a fixed tile shape for all three loops (profiling runs)
        for loop in ("density", "gradient", "force"):
            eng.set_tile_shape(loop, *shape)
    elif tune:
        tuning = eng.autotune(repeats=2)
        print(f"bench: autotuned {tuning}", file=sys.stderr, flush=True)
    run = eng.pass_graph if graph else eng.pass_batched

    def step():
        eng.restore_state()
        return run()

    for _ in range(warmup):
        step()
    eng.sync()
    times = []
    stats = None
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        stats = step()          # returns after the pass has completed
        times.append(time.perf_counter() - ts)
    eng.sync()
    total = time.perf_counter() - t0
    print(f"bench: {steps} timed steps in {total:.3f} s", file=sys.stderr,
          flush=True)
    inter = stats["n_density"] + stats["n_force"]

    # Per-loop kernel times: HIP events recorded on the engine stream around
    # every neighbour-loop kernel launch (the tile build excluded) of the
    # same number of eager passes with the same kernels (a captured graph
    # carries no per-launch events).
    eng.set_timing(True)
    t1 = time.perf_counter()
    for _ in range(steps):
        eng.restore_state()
        eager_stats = eng.pass_batched()
    eng.sync()
    eager_total = time.perf_counter() - t1
    ms, nl = eng.timing()
    eng.set_timing(False)
    assert eager_stats["n_density"] == stats["n_density"]
    kernels = {}
    for k in range(3):
        avg = ms[k] / max(nl[k], 1)
        entry = {"avg_ms": avg, "launches_per_step": nl[k] / steps,
                 "ms_per_step": ms[k] / steps,
                 "share_of_step": ms[k] / (1e3 * total)}
        if k == 0:
            # every density launch of the pass: the first round and the
            # ghost's re-runs, charged with the interactions they evaluated
            n_eval = stats["n_density_all"]
            flop = FLOP_DENSITY
        elif k == 1:
            n_eval = stats["n_gradient"]
            flop = FLOP_GRADIENT
        else:
            n_eval = stats["n_force"]
            flop = FLOP_FORCE
        entry["interactions_per_step"] = n_eval
        entry["tflops"] = n_eval * flop * steps / (ms[k] * 1e-3) / 1e12
        kernels[LOOPS[k]] = entry
    dom = int(np.argmax(ms))
    name = LOOPS[dom]
    avg_ms = kernels[name]["avg_ms"]
    achieved = kernels[name]["tflops"]
    shape = "x".join(str(c) for c in eng.get_tile_shape(name.split("_")[0]))
    prof = _profile_values(f"{name}_{shape}_{traffic_key}")
    traffic = prof.get("hbm_bytes_per_launch")
    valu = _profile_values(f"{name}_{shape}_{traffic_key}")
    roofline = {"bound": "valu", "kernel": name, "achieved": achieved,
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                "hbm_gbs": (traffic / (avg_ms * 1e-3) / 1e9) if traffic else None,
                "valu_busy": valu.get("valu_busy"),
                "valu_insts_per_interaction": valu.get("valu_insts_per_interaction"),
                "avg_launch_ms": avg_ms,
                "flop_per_interaction": [FLOP_DENSITY, FLOP_GRADIENT,
                                         FLOP_FORCE][dom]}
    out = {"workload": workload, "particles": int(st["h"].shape[0]),
           "active": int((st["time_bin"] <= props.max_active_bin).sum()),
           "value": steps * inter / total, "unit": "interactions/s",
           "ms_per_step": 1e3 * total / steps,
           "p50_step_ms": 1e3 * _pct(times, 50),
           "p99_step_ms": 1e3 * _pct(times, 99),
           "interactions_per_step": {"density": stats["n_density"],
                                     "force": stats["n_force"],
                                     "gradient": stats["n_gradient"]},
           "ghost_rounds": stats["ghost_rounds"],
           "pass": "graph replay" if graph else "eager batched",
           "eager_ms_per_step": 1e3 * eager_total / steps,
           "launches_per_step": stats["launches"],
           "tuning": tuning, "kernels": kernels, "roofline": roofline}
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
                        "launches_per_step": ps["launches"],
                        # the plain task order sums in another order, so h
                        # and borderline neighbour counts may differ slightly
                        "count_rel_diff": abs(pinter - inter) / max(inter, 1)}
        out["batched_over_plain"] = out["value"] / out["plain"]["value"]
        print(f"bench: plain pass {pdt:.2f} s", file=sys.stderr, flush=True)
    eng.close()
    return out


def _native_port():
    """The CPU baseline library built for this host (-march=native); the
    prebuilt x86-64-v3 build if the host compiler is missing."""
    path = os.path.join(ROOT, "oracle", "libcpuport_native.so")
    try:
        # always rebuilt (-B): a copy built on another host may not run here
        subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"),
                        "native"], check=True, capture_output=True, timeout=120)
        return path, "-march=native"
    except (OSError, subprocess.SubprocessError):
        return None, "-march=x86-64-v3 (prebuilt)"


def cpu_baseline(props, st, threads: int, small, runs: int = 5,
                 budget_s: float = 6.0) -> dict:
    """The oracle's FP32 restatement of the same pass on the host cores:
    one warm-up pass, then `runs` timed passes, median.  When the warm-up
    takes longer than `budget_s`, the timed passes keep a random subset of
    the active particles (the same box, fewer active) so the leg stays
    within a few tens of seconds.  The single-thread figure runs the same
    construction at a smaller size (`small` = (props, state)), median of 3."""
    from oracle import oracle as orc
    from swift_amd import ics
    port, flags = _native_port()

    def one(p, state, nthreads):
        cs = ics.copy_state(state)
        ts = time.perf_counter()
        o = orc.pass_fp32(p, cs, nthreads=nthreads, port=port or True)
        dt = time.perf_counter() - ts
        ci = int(o["n_density"].astype(np.int64).sum()) + int(
            o["n_force"].astype(np.int64).sum())
        return ci / dt, dt

    def median(p, state, nthreads, k):
        r = sorted(one(p, state, nthreads) for _ in range(k))
        return r[len(r) // 2], [v for v, _ in r]

    def n_active(p, state):
        return int((state["time_bin"] <= p.max_active_bin).sum())

    _, tw = one(props, st, threads)
    sample = st
    if tw > budget_s:
        sample = ics.copy_state(st)
        act = np.flatnonzero(sample["time_bin"] <= props.max_active_bin)
        rng = np.random.default_rng(5)
        off = rng.permutation(act)[int(round(budget_s / tw * act.size)):]
        sample["time_bin"][off] = ics.NUM_TIME_BINS + 1
    (med, med_dt), all_rates = median(props, sample, threads, runs)
    sp, ss = small
    one(sp, ss, 1)
    (med1, dt1), _ = median(sp, ss, 1, 3)
    n, nact = st["h"].shape[0], n_active(props, sample)
    n1 = ss["h"].shape[0]
    return {"value": med, "unit": "interactions/s", "cores": threads,
            "kind": "port",
            "sample": f"median of {runs} full passes (density+ghost+gradient+"
                      f"force, one warm-up before) over the headline's clustered "
                      f"{round(n ** (1 / 3))}^3 state with {nact} active particles"
                      f" ({100.0 * nact / n:.2f} %), {med_dt:.2f} s wall each on "
                      f"{threads} threads pinned one per core (the oracle's FP32 "
                      "restatement with sorted per-axis candidate lists, "
                      f"-O3 -ffast-math -funroll-loops {flags}, OpenMP over cells)",
            "runs_interactions_per_s": all_rates,
            "single_thread": {"value": med1, "unit": "interactions/s",
                              "sample": f"median of 3 passes (one warm-up before) "
                                        f"over the same construction at "
                                        f"{round(n1 ** (1 / 3))}^3 ({n_active(sp, ss)}"
                                        f" active), {dt1:.2f} s wall each on 1 "
                                        "thread"}}


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
                 chunk: int = 0, sched: bool = False) -> dict:
    """Multi-step run of the clustered construction: forces at t = 0 for all
    particles (the initial h estimate converged by the ghost), then `nsteps`
    drift-kick steps on a fixed 4-level time-bin hierarchy (bins 1..4 hold
    1.5 / 4.5 / 19 / 75 % of the particles, so 1.5, 6, 25 or 100 % are
    active), state resident in HBM, re-binned on the device when particles
    leave the stencil (a bulk flow of (150, 80, -40) carries them across
    cell faces)."""
    from swift_amd import Engine, HydroProps, ics
    from swift_amd.engine import integration
    N = n_side ** 3
    t0 = time.perf_counter()
    st = ics.clustered_box(N, seed=seed, n_spheres=1000, a_min=0.002,
                           a_max=0.02)
    ics.set_bin_hierarchy(st, {1: 0.015, 2: 0.045, 3: 0.19, 4: 0.75},
                          seed=seed + 3)
    # a bulk flow carries the particles out of their cells' stencils within
    # the run, so the device re-binning runs at this size too
    st["v"] += np.asarray((150.0, 80.0, -40.0), np.float32)
    t_ics = time.perf_counter() - t0
    print(f"bench: evolve box built ({t_ics:.1f} s)", file=sys.stderr, flush=True)
    # time line: bin 1 steps are 4 units of 3e-7; over 20 steps the bulk
    # flow moves every particle by ~0.004, a third of a 1/96 cell
    props = HydroProps(ncell=(ncell,) * 3, time_base=3e-7)
    eng = Engine(props)
    ts = time.perf_counter()
    eng.set_particles(st)
    eng.sync()
    t_up = time.perf_counter() - ts
    del st
    eng.set_integration(integration(fixed_bins=True))
    # passes through the reference's task graph, released to the device
    # as bundles, instead of the device work queue
    eng.set_step_sched(sched)
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
                    " state resident in HBM" + (
                        ", passes through the hydro task graph released as "
                        "bundles" if sched else ", passes through the device "
                        "work queue"),
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
    ap.add_argument("--n", type=int, default=256,
                    help="particles per axis of the clustered headline box "
                         "(0: skip it, e.g. to profile the other workloads; "
                         "the line then carries no value)")
    ap.add_argument("--ncell", type=int, default=48,
                    help="top-level cells per axis of the headline box")
    ap.add_argument("--modulated-n", type=int, default=128,
                    help="particles per axis of the all-active modulated box "
                         "(side line; 0: skip)")
    ap.add_argument("--modulated-ncell", type=int, default=32)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads of the CPU baseline (0: OMP_NUM_THREADS "
                         "if set, else every core this process may run on)")
    ap.add_argument("--shape", default="",
                    help="CI:CJ tile shape of all loops instead of the "
                         "autotuner's choice (e.g. 4:32)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default schedule and tile shapes")
    ap.add_argument("--graph", type=int, default=1,
                    help="1: replay each step from captured HIP graphs")
    ap.add_argument("--serving-s", type=float, default=5.0,
                    help="seconds of the serving stream (0: skip)")
    ap.add_argument("--serving-rate", type=float, default=200.0)
    ap.add_argument("--serving-sweep", default="100,300,400",
                    help="offered loads (req/s) of the latency sweep")
    ap.add_argument("--serving-sweep-s", type=float, default=3.0)
    ap.add_argument("--evolve-n", type=int, default=512,
                    help="particles per axis of the multi-step clustered run "
                         "(0: skip; 512 is the full-size case)")
    ap.add_argument("--evolve-sched-steps", type=int, default=0,
                    help="steps of the same run with the passes driven by the "
                         "task graph (0: skip)")
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

    shape = tuple(int(c) for c in args.shape.split(":")) if args.shape else None
    if args.n > 0:
        props, st, workload = clustered_state(args.n, args.ncell)
        head = measure(props, st, workload, args.steps, args.warmup,
                       not args.no_tune, bool(args.graph), not args.no_plain,
                       f"clustered{args.n}", shape)
    else:
        st = None
        workload = "none (headline skipped with --n 0)"
        head = {"value": None, "ms_per_step": None, "p50_step_ms": None,
                "p99_step_ms": None, "particles": 0, "active": 0,
                "roofline": None}
    out = {
        "metric": "SPH pair interactions/s (density+force) batched vs plain on GPU; p50/p99 step ms",
        "value": head["value"],
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "p50_step_ms": head["p50_step_ms"],
        "p99_step_ms": head["p99_step_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 positions)",
        "data": "synthetic",
        "config": {"workload": workload, "particles": head["particles"],
                   "active": head["active"], "top_cells": args.ncell ** 3,
                   "parallelism": "single GPU"},
        "roofline": head["roofline"],
    }
    for k in ("interactions_per_step", "ghost_rounds", "pass",
              "eager_ms_per_step", "launches_per_step", "tuning", "kernels",
              "plain", "batched_over_plain"):
        if k in head:
            out[k] = head[k]

    if not args.no_cpu and st is not None:
        threads = (args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", 0))
                   or len(os.sched_getaffinity(0)))
        # the single-thread sample: the same construction at 64^3
        small = clustered_state(64, 12)[:2]
        out["cpu_baseline"] = cpu_baseline(props, st, threads, small)
        print(f"bench: cpu baseline {out['cpu_baseline']['value']:.3e} "
              "interactions/s", file=sys.stderr, flush=True)
    del st

    if args.modulated_n > 0:
        mp, ms_, mw = modulated_state(args.modulated_n, args.modulated_ncell)
        out["modulated"] = measure(mp, ms_, mw, args.steps, args.warmup,
                                   not args.no_tune, bool(args.graph),
                                   not args.no_plain,
                                   f"modulated{args.modulated_n}", shape)
        del ms_
        print(f"bench: modulated {out['modulated']['ms_per_step']:.2f} ms/step",
              file=sys.stderr, flush=True)

    if args.evolve_n > 0:
        out["evolve"] = evolve_bench(args.evolve_n, args.evolve_ncell,
                                     args.evolve_steps, chunk=args.evolve_chunk)
        print(f"bench: evolve {out['evolve']['ms_per_step']:.1f} ms/step",
              file=sys.stderr, flush=True)
        if args.evolve_sched_steps > 0:
            out["evolve_sched"] = evolve_bench(
                args.evolve_n, args.evolve_ncell, args.evolve_sched_steps,
                chunk=args.evolve_chunk, sched=True)
            print(f"bench: evolve (task graph) "
                  f"{out['evolve_sched']['ms_per_step']:.1f} ms/step",
                  file=sys.stderr, flush=True)

    if args.serving_s > 0:
        out["serving"] = serving_bench(args.serving_s, args.serving_rate)
        print(f"bench: serving {out['serving']['requests']} requests",
              file=sys.stderr, flush=True)
        # latency against offered load
        sweep = []
        for rate in [float(r) for r in args.serving_sweep.split(",") if r]:
            sv = serving_bench(args.serving_sweep_s, rate, seed=1)
            sweep.append({k: sv[k] for k in (
                "requests", "p50_latency_ms", "p99_latency_ms",
                "requests_per_s", "interactions_per_s")} | {"offered_per_s": rate})
            print(f"bench: serving at {rate:g} req/s: p99 "
                  f"{sv['p99_latency_ms']:.1f} ms", file=sys.stderr, flush=True)
        out["serving"]["load_sweep"] = sweep
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
