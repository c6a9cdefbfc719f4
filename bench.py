#!/usr/bin/env python3
"""Benchmark: particle-substeps/s of the XPBD cloth solver on one MI355X.

Default workload (``--workload batch512``): 512 independent cloths, sides log-uniform 16..128
(seeded), each with its own stiffness, spacing, 0-4 attached corners and 1-3 sphere colliders,
5 substeps x 10 iterations, self-collision on. Batched = every cloth a segment of one solver,
one launch per stage; before warm-up the launch-variant tuner picks graph replay, neighbour
search order, attachment load batch and XCD remap (all bit-identical). Plain = one solver per
cloth, each launching its own kernels per stage on one shared stream, no graph.

A step is one 1/60 s frame of every cloth. ``value`` = sum(particles) x substeps x steps /
seconds of the timed region (inputs resident in HBM); the timed frames are the last ones of the
workload's 120-frame horizon, after the cloths have settled on their colliders. The cloths opt
in to the tuned hash grid (``--reference-grid`` keeps the reference's). The line also carries
the plain path's rate, per-frame latency percentiles, the dominant kernel's roofline (HIP events
on the solver's stream; algorithmic bytes with dense long-range attachments at their 4-B
distance, beside the reference-layout figure, and the copy-calibrated counter bytes from
``profiles/``), the CPU baseline (the restated reference CPU solver on host cores), the serving
sweep, and the same measurements for the single-cloth configurations 1-3 (``configs``).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from velvet_amd import scenes as S  # noqa: E402
from velvet_amd.solver import ClothSolver  # noqa: E402

METRIC = "particle-substeps/s, batched vs plain; p50/p99 step latency at fixed QPS"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def hip_stream():
    hip = ctypes.CDLL("libamdhip64.so")
    s = ctypes.c_void_p()
    rc = hip.hipStreamCreateWithFlags(ctypes.byref(s), ctypes.c_uint(1))
    if rc != 0:
        raise RuntimeError(f"hipStreamCreateWithFlags failed: {rc}")
    return s.value


def copy_probe(mib):
    """One device-to-device copy of known size: the calibration kernel for HBM counters."""
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def workload_scenes(name):
    if name == "batch512":
        return S.config4(count=512), "512 cloths, sides log-uniform 16..128, 5x10, per-cloth hash grids"
    if name == "config1":
        return [S.config1()], "40x40 cloth, 2 pinned corners, sphere, 5x10, omega 1.5"
    if name == "config2":
        return [S.config2()], "256x256 cloth, 256-slot LRA edge, sphere + plane, 2^18 cells, cell = d, 5x10"
    if name == "config3":
        return [S.config3()], "1024x1024 cloth, 3 seeded spheres, 2^21 cells, 4x8"
    if name == "serve64":
        return [sc for sc, _ in serve_pool(0)], "the 64 serving job descriptions as one batch, 5x10"
    raise SystemExit(f"unknown workload {name}")


def counts(scenes):
    tot = dict(N=0, S=0, B=0, A=0, L=0, T=0)
    for sc in scenes:
        c = sc.counts()
        for k in tot:
            tot[k] += c[k]
    return tot


def time_frames(solver, frames):
    solver.synchronize()
    t0 = time.perf_counter()
    for _ in range(frames):
        solver.Simulate()
    solver.synchronize()
    return time.perf_counter() - t0


def latency(solver, frames):
    out = []
    for _ in range(frames):
        t0 = time.perf_counter()
        solver.Simulate()
        solver.synchronize()
        out.append((time.perf_counter() - t0) * 1e3)
    return out


def kernel_times(solver, frames):
    solver.set_timing(True)
    for _ in range(frames):
        solver.Simulate()
    solver.synchronize()
    rep = json.loads(solver.timing())
    solver.set_timing(False)
    return rep


def plain_rate(scenes, frames, stream):
    solvers = []
    for sc in scenes:
        s = ClothSolver(use_graph=False)
        s.load_scene(sc)
        s.set_stream(stream)
        solvers.append(s)
    for s in solvers:       # warm-up frame (commit + first launches)
        s.Simulate()
    solvers[0].synchronize()
    t0 = time.perf_counter()
    for _ in range(frames):
        for s in solvers:
            s.Simulate()
    solvers[0].synchronize()
    dt = time.perf_counter() - t0
    for s in solvers:
        s.close()
    return dt


def measured_traffic(workload, kernel="vtx::k_iterate_grid<0,"):
    """HBM bytes per launch of a kernel (the plain Jacobi pass by default) from the committed,
    copy-calibrated FETCH_SIZE / WRITE_SIZE summary of this workload (separate rocprofv3 runs):
    the fetch and write columns of the kernel's row, or None when the file, the row or either
    column is missing."""
    path = os.path.join(ROOT, "profiles", f"{workload}_hbm_counters.md")
    try:
        cols = None
        for ln in open(path):
            if not ln.startswith("|"):
                continue
            cells = [c.strip() for c in ln.strip().strip("|").split("|")]
            if cells[0] == "kernel":
                cols = ([k for k, c in enumerate(cells) if c.startswith("fetch MB")] +
                        [k for k, c in enumerate(cells) if c.startswith("write MB")])
            elif cols and len(cols) == 2 and cells[0].startswith(f"`void {kernel}"):
                return (float(cells[cols[0]]) + float(cells[cols[1]])) * 1e6, os.path.relpath(path, ROOT)
    except (OSError, ValueError):
        pass
    return None, None


def jacobi_bytes(scenes):
    """Algorithmic bytes of one Jacobi pass over the workload: the reference-layout figure
    12S + 20B + 12A + 12L + 28N, except that a cloth whose attachments are the dense long-range
    form (every particle to every slot, in slot order) needs only the 4-B distance per
    attachment (particle and slot follow from the position in the table)."""
    tot = 0
    for sc in scenes:
        c = sc.counts()
        dense = all(cl.attach_pid.shape[0] > 0 and cl.attach_pid.shape[0] == cl.slots.shape[0] * cl.n
                    for cl in sc.cloths) and c["A"] > 0
        tot += 12 * c["S"] + 20 * c["B"] + (4 if dense else 12) * c["A"] + 12 * c["L"] + 28 * c["N"]
    return tot


def dominant(kt):
    """(label, entry) of the launch group with the largest total time in a timing report."""
    items = [(k, v) for k, v in kt.items() if k != "total" and v.get("launches")]
    return max(items, key=lambda kv: kv[1]["ms"])


def config_leg(name, steps, warmup, tune_frames, cpu_seconds):
    """One of the single-cloth configurations: tuned, settled so that the timed frames end at
    the configuration's horizon, timed, per-stage times, the dominant kernel's roofline and a
    bounded sample of the restated reference CPU solver."""
    scenes, desc = workload_scenes(name)
    horizon = scenes[0].frames
    tot = counts(scenes)
    substeps = scenes[0].params.num_substeps
    for sc in scenes:
        sc.params.tune_hash_grid = 1
    solver = ClothSolver(use_graph=True)
    for sc in scenes:
        solver.load_scene(sc)
    tuning = solver.autotune(frames=tune_frames) if tune_frames > 0 else solver.tuning()
    settle = max(0, horizon - warmup - steps)
    for _ in range(settle + warmup):
        solver.Simulate()
    dt = time_frames(solver, steps)
    rate = tot["N"] * substeps * steps / dt
    kt = kernel_times(solver, 3)
    solver.close()
    label, ent = dominant(kt)
    avg_ms = ent["ms"] / ent["launches"]
    out = {"workload": f"{name}: {desc}", "particles": tot["N"], "value": rate, "unit": "particle-substeps/s",
           "ms_per_step": dt / steps * 1e3, "frames_timed": [settle + warmup + 1, settle + warmup + steps],
           "tuning": tuning, "dominant": {"label": label, "kernel": ent["kernel"], "avg_launch_ms": round(avg_ms, 6),
                                          "launches_per_frame": ent["launches"] / 3},
           "kernels_ms": kt}
    if ent["kernel"] == "k_iterate":
        alg = jacobi_bytes(scenes)
        ach = alg / (avg_ms * 1e-3) / 1e9
        traffic, src = measured_traffic(name)
        out["roofline"] = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": src,
                           "algorithmic_bytes_per_launch": alg,
                           "reference_layout_bytes_per_launch": S.algorithmic_bytes_per_iteration(tot)}
        if traffic:
            out["roofline"]["traffic_frac"] = round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    else:
        out["roofline"] = {"bound": "launch", "note": f"{ent['launches'] / 3:.0f} launches of {ent['kernel']} per "
                                                      f"frame average {avg_ms * 1e3:.1f} us; the frame is launch "
                                                      f"bound, not HBM bound"}
    if cpu_seconds > 0:
        rate_c, spent, frames, _ = cpu_baseline(scenes, cpu_seconds, warm=tot["N"] < 500_000)
        out["cpu_baseline"] = {"value": rate_c, "unit": "particle-substeps/s", "cores": 1, "kind": "port",
                               "sample": f"restated reference CPU solver, single thread, {frames} timed frame(s) of "
                                         f"this configuration, {spent:.1f} s"}
    return out


def cpu_baseline(scenes, budget_s, warm=True):
    """The restated reference CPU solver on one thread: one timed frame per cloth (after a
    warm-up frame unless warm is False, for cloths whose single frame fills the budget), cloths
    in turn until the budget is spent."""
    import oracle
    done_ps, spent, frames, cloths = 0, 0.0, 0, 0
    k = 0
    while spent < budget_s and k < 4 * len(scenes):
        sc = scenes[k % len(scenes)]
        cpu = oracle.CpuSolver(sc)
        if warm:
            cpu.simulate()                 # warm-up frame, excluded
        t0 = time.perf_counter()
        cpu.simulate()
        dt = time.perf_counter() - t0
        spent += dt
        done_ps += sc.n * sc.params.num_substeps
        frames += 1
        cloths += 1
        k += 1
        del cpu
    return done_ps / spent, spent, frames, cloths


def cpu_baseline_threads(scenes, budget_s, workers):
    """The same restated reference CPU solver on `workers` host threads at once, each taking the
    workload's cloths in turn (the cloths are independent; the solver calls release the GIL).
    Per thread: one warm-up frame per cloth, then one timed frame; the aggregate rate is the sum
    of the threads' rates over their timed frames, which all run concurrently."""
    import threading
    import oracle
    lock = threading.Lock()
    state = {"next": 0}
    rates = [0.0] * workers
    counts = [0] * workers
    deadline = time.perf_counter() + budget_s

    def work(w):
        done_ps, spent = 0, 0.0
        while time.perf_counter() < deadline:
            with lock:
                k = state["next"]
                state["next"] += 1
            sc = scenes[k % len(scenes)]
            cpu = oracle.CpuSolver(sc)
            cpu.simulate()
            t0 = time.perf_counter()
            cpu.simulate()
            spent += time.perf_counter() - t0
            done_ps += sc.n * sc.params.num_substeps
            counts[w] += 1
            del cpu
        rates[w] = done_ps / spent if spent else 0.0

    threads = [threading.Thread(target=work, args=(w,)) for w in range(workers)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    return sum(rates), sum(counts)


def parse_tuning(text):
    """'k=v,k=v' -> {k: int(v)}"""
    return {k: int(v) for k, v in (kv.split("=") for kv in text.split(",") if kv)}


def serve_pool(seed, count=64):
    """Job descriptions (scene, horizon) the serving run's requests cycle through: sizes
    log-uniform 32^2..512^2, horizons U(30, 600), attachments and colliders as drawn per job.
    Building a cloth is the client's work, so it happens before the timed run."""
    pool = [S.serving_job(seed, k) for k in range(count)]
    for sc, _ in pool:
        sc.params.tune_hash_grid = 1       # the server's tuning picks the grid (hash_grid)
    return pool


def serve_run(rate, seconds, pool, max_live_particles, use_graph, step_budget_ms=0.0, ahead=4,
              reserve_particles=6_000_000, tuning=None, pipelined=True):
    """Poisson arrivals at `rate` req/s for `seconds` of wall time against the serving loop.

    Request k is job description pool[k % len(pool)]. Arrived requests wait in a backlog; the
    loop submits them (the server packs each on the host) while the server's own queue holds
    fewer than `ahead` jobs, so host memory stays bounded when arrivals exceed capacity. Step
    latency is the server's own wall time per step;
    admission wait is the time from a request's arrival to the end of the step that admitted it
    (requests still waiting at the end are counted in `backlog`, not in the wait percentiles).
    Finished jobs are read back by the server and released here. `backlog_growth_per_s` is the
    least-squares slope of the backlog (arrived, not yet admitted) over the second half of the
    window: near zero when the server keeps up with the arrivals.
    """
    from velvet_amd.server import Server
    rng = np.random.default_rng(0)
    lat, work, live, waits = [], 0, [], []
    phases = {"admit": [], "frame": [], "retire": []}
    backlog = []
    trace_t, trace_b = [], []
    arrival = {}
    arrived = submitted = admitted = completed = 0
    k = 0
    next_t = rng.exponential(1.0 / rate)
    with Server(max_live_particles=max_live_particles, use_graph=use_graph,
                step_budget_ms=step_budget_ms, reserve_particles=reserve_particles, pipelined=pipelined) as sv:
        if tuning:
            sv.solver.set_tuning(**tuning)
        tun = sv.solver.tuning()
        queued = 0
        waiting, running = {}, set()
        t0 = time.perf_counter()
        while True:
            now = time.perf_counter() - t0
            if now >= seconds:
                break
            while next_t <= now:
                backlog.append(k)
                arrival[k] = next_t
                k += 1
                arrived += 1
                next_t += rng.exponential(1.0 / rate)
            while backlog and queued < ahead:
                r = backlog.pop(0)
                sc, horizon = pool[r % len(pool)]
                waiting[sv.submit(sc, horizon)] = arrival.pop(r)
                submitted += 1
                queued += 1
            st = sv.step()
            t_end = time.perf_counter() - t0
            queued = st.queued_jobs
            trace_t.append(t_end)
            trace_b.append(len(backlog) + queued)
            admitted += st.admitted
            completed += st.completed
            if st.admitted:
                for j in [j for j in waiting if sv.status(j)[0] in (2, 3, 4)]:   # running, done, failed
                    waits.append((t_end - waiting.pop(j)) * 1e3)
                    running.add(j)
            if st.completed:
                for j in [j for j in running if sv.status(j)[0] in (3, 4)]:
                    sv.release(j)
                    running.discard(j)
            if st.live_jobs or st.completed:
                lat.append(st.step_ms)
                phases["admit"].append(st.admit_ms)
                phases["frame"].append(st.frame_ms)
                phases["retire"].append(st.retire_ms)
                work += st.particle_substeps
                live.append(st.live_particles)
        wall = time.perf_counter() - t0
    pct = lambda a, q: float(np.percentile(a, q)) if a else None
    half = [i for i, t in enumerate(trace_t) if t >= wall / 2]
    growth = float(np.polyfit(np.asarray(trace_t)[half], np.asarray(trace_b)[half], 1)[0]) if len(half) > 2 else None
    return {"rate_rps": rate, "seconds": round(wall, 2), "steps": len(lat),
            "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
            "particle_substeps_per_s": work / wall, "arrived": arrived, "admitted": admitted,
            "completed": completed, "backlog": len(backlog) + queued, "backlog_growth_per_s": growth,
            "distinct_jobs": len(pool),
            "admit_wait_p50_ms": pct(waits, 50), "admit_wait_p99_ms": pct(waits, 99),
            "phase_p50_ms": {k: pct(v, 50) for k, v in phases.items()},
            "phase_p99_ms": {k: pct(v, 99) for k, v in phases.items()},
            "mean_live_particles": float(np.mean(live)) if live else 0.0,
            "max_live_particles_cap": max_live_particles, "step_budget_ms": step_budget_ms,
            "reserve_particles": reserve_particles, "pipelined": pipelined, "tuning": tun}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="batch512")
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--latency-steps", type=int, default=30)
    ap.add_argument("--kernel-steps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--tune-frames", type=int, default=16,
                    help="frames per variant for the launch-variant tuner before warm-up; 0 = defaults")
    ap.add_argument("--set-tuning", default="", metavar="K=V,...",
                    help="fix launch variants instead of tuning, e.g. use_graph=1,neighbor_order=1")
    ap.add_argument("--reference-grid", action="store_true",
                    help="keep every cloth on the reference hash grid (tune_hash_grid = 0); by default "
                         "the cloths opt in and the tuner picks the hash-grid resolution")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-rates", default="50,60,100,200,400", help="comma-separated Poisson rates (req/s)")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-cap", type=int, default=0, help="live-particle cap; 0 = none")
    ap.add_argument("--serve-budget-ms", type=float, default=11.0,
                    help="admission budget on the predicted frame time")
    ap.add_argument("--serve-graph", type=int, default=0)
    ap.add_argument("--serve-pipelined", type=int, default=1,
                    help="1: each serving step returns once its frame is launched, so request handling "
                         "overlaps the frame; 0: each step waits for its frame")
    ap.add_argument("--serve-tuning", default="neighbor_order=2,hash_lds=0,streams=2,hash_grid=1,lra_batch=4", metavar="K=V,...",
                    help="launch variants of the serving solver (the live set changes every step, so it is "
                         "not tuned online); the default is the measured best of a sweep on this job mix")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--configs", default="config1,config2,config3",
                    help="single-cloth configurations measured beside the workload ('' = none)")
    ap.add_argument("--config-steps", type=int, default=20)
    ap.add_argument("--config-cpu-seconds", type=float, default=4.0)
    ap.add_argument("--serve-long-rates", default="50,60,100",
                    help="rates served for --serve-long-seconds instead of --serve-seconds")
    ap.add_argument("--serve-long-seconds", type=float, default=20.0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1 (batched and "
                                   "served jobs would shard across GPUs with no exchange)"}))
        sys.exit(2)

    scenes, desc = workload_scenes(args.workload)
    for sc in scenes:
        sc.params.tune_hash_grid = 0 if args.reference_grid else 1
    horizon = scenes[0].frames
    tot = counts(scenes)
    substeps = scenes[0].params.num_substeps
    iters = scenes[0].params.num_iterations
    log(f"workload {args.workload}: {len(scenes)} cloths, {tot}")

    solver = ClothSolver(use_graph=True)
    for sc in scenes:
        solver.load_scene(sc)
    tuning = solver.tuning()
    if args.set_tuning:
        solver.set_tuning(**parse_tuning(args.set_tuning))
        tuning = solver.tuning()
    elif args.tune_frames > 0:
        log(f"tuning launch variants ({args.tune_frames} frames each)")
        tuning = solver.autotune(frames=args.tune_frames)
        log(f"  {tuning}")
    # the timed frames end at the workload's horizon (120 frames for the batch): the cloths have
    # fallen onto their colliders and rest there, with the contacts and neighbour lists of the
    # draped state
    settle = max(0, horizon - args.warmup - args.steps)
    log(f"registered; {settle} settling + {args.warmup} warm-up frames")
    for _ in range(settle + args.warmup):
        solver.Simulate()
    solver.synchronize()
    log(f"timing {args.steps} frames")
    dt = time_frames(solver, args.steps)
    value = tot["N"] * substeps * args.steps / dt
    ms_per_step = dt / args.steps * 1e3
    log(f"batched: {value:.4e} particle-substeps/s, {ms_per_step:.3f} ms/frame")

    lat = latency(solver, args.latency_steps)
    kt = kernel_times(solver, args.kernel_steps)
    it_ms = sum(v["ms"] for v in kt.values() if v.get("kernel") == "k_iterate")
    it_n = sum(v["launches"] for v in kt.values() if v.get("kernel") == "k_iterate")
    avg_ms = it_ms / it_n if it_n else float("nan")
    alg_bytes = jacobi_bytes(scenes)
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(args.workload)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": traffic_src, "kernel": "k_iterate", "avg_launch_ms": round(avg_ms, 6),
                "algorithmic_bytes_per_launch": alg_bytes,
                "reference_layout_bytes_per_launch": S.algorithmic_bytes_per_iteration(tot)}
    if traffic:
        roofline["traffic_frac"] = round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    solver.close()

    if args.copy_probe:
        copy_probe(args.copy_probe)

    plain = None
    if not args.no_plain and len(scenes) > 1:
        log(f"plain path: {len(scenes)} solvers, {args.plain_steps} frames")
        pdt = plain_rate(scenes, args.plain_steps, hip_stream())
        pval = tot["N"] * substeps * args.plain_steps / pdt
        plain = {"value": pval, "ms_per_step": pdt / args.plain_steps * 1e3,
                 "batched_over_plain": value / pval}
        log(f"plain: {pval:.4e} particle-substeps/s -> batched/plain {value / pval:.2f}x")

    serving = serving_sustained = None
    if not args.no_serve:
        serving = []
        log("serving: building the job descriptions")
        pool = serve_pool(0)
        long_rates = {float(x) for x in args.serve_long_rates.split(",") if x}
        for r in [float(x) for x in args.serve_rates.split(",") if x]:
            secs = args.serve_long_seconds if r in long_rates else args.serve_seconds
            log(f"serving: {r:g} req/s for {secs:g} s, cap {args.serve_cap} live particles")
            serving.append(serve_run(r, secs, pool, args.serve_cap, bool(args.serve_graph),
                                     args.serve_budget_ms, tuning=parse_tuning(args.serve_tuning),
                                     pipelined=bool(args.serve_pipelined)))
            log(f"  {serving[-1]}")
        # the highest swept rate whose backlog stopped growing (second-half slope under 2% of
        # the arrival rate) over a long window
        ok = [r["rate_rps"] for r in serving if r["rate_rps"] in long_rates and r["backlog_growth_per_s"] is not None
              and r["backlog_growth_per_s"] <= 0.02 * r["rate_rps"]]
        serving_sustained = max(ok) if ok else None

    configs = {}
    for name in [x for x in args.configs.split(",") if x]:
        log(f"{name}: tuned, {args.config_steps} frames timed at the end of its horizon")
        configs[name] = config_leg(name, args.config_steps, 2, 4 if args.tune_frames > 0 else 0,
                                   0.0 if args.no_cpu else args.config_cpu_seconds)
        log(f"  {configs[name]['value']:.4e} particle-substeps/s, {configs[name]['ms_per_step']:.3f} ms/frame")

    cpu = cpu_threads = None
    if not args.no_cpu:
        log(f"cpu baseline ({args.cpu_seconds:.0f} s budget)")
        rate, spent, frames, cloths = cpu_baseline(scenes, args.cpu_seconds)
        cpu = {"value": rate, "unit": "particle-substeps/s", "cores": 1, "kind": "port",
               "sample": f"restated reference CPU solver (Gauss-Seidel, VtClothSolverCPU + "
                         f"SpatialHashCPU), single thread, 1 timed frame each of the first {cloths} "
                         f"cloths of the workload after a warm-up frame, {spent:.1f} s",
               "host_cpu": _cpu_model()}
        workers = min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {workers} threads ({args.cpu_seconds:.0f} s budget)")
        mrate, mcloths = cpu_baseline_threads(scenes, args.cpu_seconds, workers)
        cpu_threads = {"value": mrate, "unit": "particle-substeps/s", "cores": workers, "kind": "port",
                       "sample": f"the same solver on {workers} host threads at once, one timed frame "
                                 f"per cloth after a warm-up frame, {mcloths} cloths, sum of the "
                                 f"threads' rates"}

    line = {"metric": METRIC, "value": value, "unit": "particle-substeps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded cloths)",
            "config": {"workload": f"{args.workload}: {desc}", "particles": tot["N"],
                       "stretch": tot["S"], "bend": tot["B"], "attach": tot["A"],
                       "substeps": substeps, "iterations": iters, "tuning": tuning,
                       "frames_timed": [settle + args.warmup + 1, settle + args.warmup + args.steps]},
            "plain": plain,
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                           "frames": len(lat)},
            "kernels_ms": {k: v for k, v in kt.items()},
            "serving": serving, "serving_sustained_rate_rps": serving_sustained, "roofline": roofline, "cpu_baseline": cpu,
            "cpu_baseline_threads": cpu_threads, "configs": configs}
    print(json.dumps(line), flush=True)


def _cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip() + f" ({os.cpu_count()} logical)"
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
