#!/usr/bin/env python3
"""Benchmark of the Voxelyze3 physics path on one MI355X.

Workload: 512 uneven robots on a 12^3 grid (Bernoulli(0.5) occupancy, largest
face-connected component, four materials, seed 0), with the physics of the
single-cube configuration (gravity, floor with static/kinetic friction, 2 Hz
phase actuation with CTE 0.01, Poisson effects).  One "step" is one physics
time step of every robot of the batch; the timed region holds exactly --steps
steps after --warmup untimed ones, with the batch's stream synchronized on both
sides, state resident in HBM throughout.

The headline runs with FP64 state (--precision, default f64): on these robots,
friction stick/slip amplifies rounding differences, and only FP64 state keeps
every robot's fitness within 0.5% of the CPU engine over the 20,000-step
horizon.  The FP32 state path is timed beside it (`f32`).

Unless --block/--steps-per-launch are given, the batch configuration (preferred
workgroup size, steps per resident launch) is the autotuner's choice
(vx3h_autotune) on this same workload, reported under `autotune`.

Printed: one JSON line with voxel-steps/s (value), the roofline of the dominant
kernel (HIP events on the engine's stream), the plain one-simulation-per-launch
HIP path on the same robots, and the CPU engine timed on this host's cores on a
bounded sample of whole robots, whose final fitness values also check the GPU's.
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_step(V, L, f64):
    """Compulsory structure-of-arrays bytes of one step, each array touched once.
    Voxel: position, linear momentum, orientation, angular momentum read+written,
    flags read+written, phase, material index, link-table entry.  Link: pos2,
    angle1v, angle2v, strain, maxStrain, strainOffset, flags read+written,
    material index, two endpoints, two link-table entries."""
    if f64:
        return 232 * V + 196 * L   # 13*8*2 + 8 + 8 + 4 + 4 ; (9*8 + 16)*2 + 4 + 8 + 8
    return 124 * V + 124 * L       # 13*4*2 + 8 + 4 + 4 + 4 ; 13*4*2 + 4 + 8 + 8


def run_batch(models, precision, path, block, steps, warmup, kt_steps, spl):
    """Times `steps` steps.  Kernel durations: HIP events around every resident launch
    of the timed region (mean per launch of `spl` steps), and -- for batches with
    stream-path simulations -- per-pass events over `kt_steps` extra steps."""
    from gpuvoxels_amd import _abi, engine
    b = engine.Batch(precision=precision, path=path, use_graph=1, block_size=block, steps_per_launch=spl)
    for m in models:
        b.admit(m)
    b.step(warmup)
    b.sync()
    b.set_timing(True)
    b.timing()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    dt = time.perf_counter() - t0
    res_ms, res_n, res_steps = b.timing()
    b.set_timing(False)
    st = b.poll()
    all_running = all(s.state == _abi.SIM_RUNNING for s in st.values())
    kt = {}
    if res_n:
        kt["resident"] = res_ms / res_n
    if kt_steps and not res_n:
        for k, v in b.kernel_times(kt_steps).items():
            if k != "resident":
                kt[k] = v
    b.close()
    return dt, kt, all_running, (res_n, res_steps)


def plain_path(models, precision, steps):
    """One simulation per launch sequence: every robot in its own single-sim batch on
    the stream path, stepped one step at a time in turn from one host thread."""
    from gpuvoxels_amd import _abi, engine
    bs = []
    for m in models:
        b = engine.Batch(precision=precision, path=_abi.PATH_STREAM, use_graph=0)
        b.admit(m)
        bs.append(b)
    for b in bs:
        b.step(2)
    for b in bs:
        b.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        for b in bs:
            b.step(1)
    for b in bs:
        b.sync()
    dt = time.perf_counter() - t0
    vs = sum(m.n_voxels for m in models) * steps / dt
    for b in bs:
        b.close()
    return vs, dt


def cpu_baseline(models, steps, threads, kind):
    """The CPU engine on whole robots for `steps` steps, one robot per thread at a
    time.  kind 'ref' is the reference's own CPU sources compiled in-tree, 'oracle'
    the CPU restatement.  Returns (voxel-steps/s, seconds, fitness per robot)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OSim, load
    load(kind)
    work = list(enumerate(models))
    lock = threading.Lock()
    done = [0]
    fit = [None] * len(models)

    def worker():
        while True:
            with lock:
                if not work:
                    return
                i, m = work.pop()
            o = OSim(m, kind=kind)
            o.step(steps)
            f = o.status().fitness
            with lock:
                done[0] += m.n_voxels * steps
                fit[i] = f

    ts = [threading.Thread(target=worker) for _ in range(threads)]
    t0 = time.perf_counter()
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    dt = time.perf_counter() - t0
    return done[0] / dt, dt, fit


def gpu_fitness(models, precision, steps):
    from gpuvoxels_amd import _abi, engine
    b = engine.Batch(precision=precision, path=_abi.PATH_AUTO)
    slots = [b.admit(m) for m in models]
    b.step(steps)
    st = b.poll()
    out = [st[s].fitness for s in slots]
    b.close()
    return out


def serving_run(precision, seconds, rate, seed=4):
    """Poisson arrivals of robots whose voxel counts are log-uniform in [100, 50000]
    (the 12^3 robot generator on scaled grids) and horizons uniform in [2000, 20000]
    steps, served by the continuous-batching loop; per-request latency is submit to
    result.  Returns (summary dict, request models)."""
    import numpy as np
    from gpuvoxels_amd import engine, workloads
    rng = np.random.default_rng(seed)
    t, reqs = 0.0, []
    while t < seconds:
        t += rng.exponential(1.0 / rate)
        target = float(np.exp(rng.uniform(np.log(100), np.log(50000))))
        grid = max(5, int(round((target / 0.475) ** (1.0 / 3.0))))
        m = workloads.robot(rng, grid=grid, steps=0, name="req")
        reqs.append((t, m, int(rng.integers(2000, 20001))))
    sv = engine.Server(precision=precision, quantum_steps=500)
    lat, vsteps, i, tickets = [], 0.0, 0, {}
    t0 = sv.clock()
    first = last = None
    while True:
        now = sv.clock() - t0
        while i < len(reqs) and reqs[i][0] <= now:
            tickets[sv.submit(reqs[i][1], reqs[i][2])] = i
            i += 1
        active = sv.round()
        for r in sv.results():
            lat.append(r.t_done - r.t_submit)
            vsteps += r.voxel_steps
            first = r.t_submit if first is None else min(first, r.t_submit)
            last = r.t_done if last is None else max(last, r.t_done)
        if i == len(reqs) and active == 0:
            break
        if active == 0 and i < len(reqs):
            time.sleep(max(0.0, min(0.01, reqs[i][0] - (sv.clock() - t0))))
    sv.close()
    lat = np.array(lat)
    span = last - first
    return {"requests": len(reqs), "rate_per_s": rate, "arrival_seconds": seconds,
            "voxels_min_max": [int(min(m.n_voxels for _, m, _ in reqs)), int(max(m.n_voxels for _, m, _ in reqs))],
            "value": vsteps / span, "unit": "voxel-steps/s", "seconds": span,
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "dtype": "f64" if precision == 1 else "f32"}, reqs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--robots", type=int, default=512)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", choices=["f32", "f64"], default="f64")
    ap.add_argument("--path", choices=["auto", "stream", "resident"], default="auto")
    ap.add_argument("--block", type=int, default=None,
                    help="preferred workgroup size (0 = engine order); default: chosen by the autotuner")
    ap.add_argument("--steps-per-launch", type=int, default=None,
                    help="steps of one resident kernel launch; default: chosen by the autotuner")
    ap.add_argument("--no-autotune", action="store_true", help="block 0 and 1000 steps per launch unless given")
    ap.add_argument("--tune-steps", type=int, default=2000, help="timed steps per autotuner candidate")
    ap.add_argument("--tune-warmup", type=int, default=4000,
                    help="untimed steps per autotuner candidate (the robots' motion, and with it the share of "
                         "large-angle links, builds up over the first few thousand steps)")
    ap.add_argument("--kernel-steps", type=int, default=200, help="extra per-pass timed steps (stream-path sims only)")
    ap.add_argument("--f32-steps", type=int, default=20000, help="timed steps of the FP32 side run (0 = skip)")
    ap.add_argument("--plain-steps", type=int, default=50)
    ap.add_argument("--cpu-robots", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=20000)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serving-seconds", type=float, default=6.0, help="arrival window of the serving run (0 = skip)")
    ap.add_argument("--serving-rate", type=float, default=40.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (two launches: warm-up + timed), a kernel of "
                         "known HBM traffic for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE readings")
    ap.add_argument("--traffic", type=float, default=None,
                    help="measured HBM bytes per launch of the dominant kernel, reported as roofline.traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: simulations are independent and would be bin-packed "
              "across GPUs with no communication; this build runs and reports a single MI355X. Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    from gpuvoxels_amd import _abi, engine, workloads
    if engine.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    copy_gbs = None
    if args.calib_copy_mib:
        nb = args.calib_copy_mib << 20
        ms = engine.calib_copy(nb, reps=1)
        copy_gbs = 2 * nb / (ms * 1e-3) / 1e9
        log("plain copy of %d MiB: %.3f ms, %.0f GB/s read+write" % (args.calib_copy_mib, ms, copy_gbs))
    models = workloads.robots(args.robots, seed=args.seed, steps=0)
    V = sum(m.n_voxels for m in models)
    L = int(sum(len(engine.topology(m)[0]) for m in models))
    log("robots %d voxels %d links %d" % (len(models), V, L))
    path = {"auto": _abi.PATH_AUTO, "stream": _abi.PATH_STREAM, "resident": _abi.PATH_RESIDENT}[args.path]
    f64 = args.precision == "f64"
    prec = _abi.PREC_F64 if f64 else _abi.PREC_F32

    def tuned(p):
        """(block, steps per launch, autotuner record) for precision p: the given values,
        else the autotuner's choice on this workload (vx3h_autotune), else the defaults."""
        if (args.block is not None and args.steps_per_launch is not None) or args.no_autotune:
            return (args.block or 0), (args.steps_per_launch or 1000), None
        t0 = time.perf_counter()
        cfg, trials = engine.autotune(models, precision=p, path=path, warmup_steps=args.tune_warmup,
                                      trial_steps=args.tune_steps)
        rec = {"seconds": time.perf_counter() - t0, "trial_steps": args.tune_steps, "warmup_steps": args.tune_warmup,
               "chosen": {"block_size": cfg["block_size"], "steps_per_launch": cfg["steps_per_launch"]},
               "trials": [{"block_size": b_, "steps_per_launch": s_, "voxel_steps_per_s": v_}
                          for b_, s_, _, v_ in trials]}
        log("autotune (%s): %s in %.1f s" % ("f64" if p == _abi.PREC_F64 else "f32", rec["chosen"], rec["seconds"]))
        block = args.block if args.block is not None else cfg["block_size"]
        spl_ = args.steps_per_launch if args.steps_per_launch is not None else cfg["steps_per_launch"]
        return block, spl_, rec

    block, spl_cfg, tune_rec = tuned(prec)
    dt, kt, all_running, (res_n, res_steps) = run_batch(models, prec, path, block, args.steps, args.warmup,
                                                         args.kernel_steps, spl_cfg)
    value = V * args.steps / dt
    log("%s: timed %d steps in %.3f s -> %.3e voxel-steps/s; kernel ms %s" % (args.precision, args.steps, dt, value, kt))
    dom = max(kt, key=kt.get)
    spl = res_steps // res_n if dom == "resident" else 1
    per_step = bytes_per_step(V, L, f64)
    if dom in ("resident", "links", "voxels"):
        ab = per_step * spl if dom == "resident" else (
            (196 if f64 else 124) * L if dom == "links" else (232 if f64 else 124) * V)
        achieved = ab / (kt[dom] * 1e-3) / 1e9
    else:
        ab, achieved = None, None
    traffic = args.traffic
    if traffic is None and dom == "resident":
        # measured HBM bytes per resident launch, kept with the rocprofv3 summaries:
        # lines "<dtype> <steps per launch> <bytes>"
        tf = os.path.join(ROOT, "profiles", "hbm_traffic_resident_512robots.txt")
        if os.path.exists(tf):
            for line in open(tf):
                f = line.split()
                if len(f) == 3 and f[0] == args.precision and f[1].isdigit() and int(f[1]) == spl:
                    traffic = float(f[2])
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
            "kernel": dom, "alg_bytes_per_launch": ab, "steps_per_launch": spl, "avg_launch_ms": kt[dom],
            "alg_bytes_per_step": per_step, "kernel_ms": kt}

    out = {
        "metric": "voxel-steps/s on 512-robot batch vs plain path; p50/p99 per-sim latency",
        "value": value,
        "unit": "voxel-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {"workload": "%d uneven robots, 12^3 grid, Bernoulli(0.5) occupancy, largest face-connected "
                               "component, 4 materials, seed %d, gravity+floor friction+2 Hz actuation+Poisson, "
                               "20,000-step horizon" % (len(models), args.seed),
                   "robots": len(models), "voxels": V, "links": L, "path": args.path,
                   "block_size": block, "steps_per_launch": spl_cfg, "all_running": all_running},
        "roofline": roof,
    }
    if tune_rec:
        out["autotune"] = tune_rec
    if args.f32_steps and f64:
        block32, spl32, _ = tuned(_abi.PREC_F32)
        dt32, kt32, _, _ = run_batch(models, _abi.PREC_F32, path, block32, args.f32_steps, args.warmup, 0, spl32)
        out["f32"] = {"value": V * args.f32_steps / dt32, "unit": "voxel-steps/s", "steps": args.f32_steps,
                      "ms_per_step": dt32 / args.f32_steps * 1e3, "block_size": block32, "steps_per_launch": spl32}
        log("f32: %.3e voxel-steps/s" % out["f32"]["value"])
    if not args.no_plain:
        pv, pdt = plain_path(models, prec, args.plain_steps)
        out["plain_path"] = {"value": pv, "unit": "voxel-steps/s", "steps": args.plain_steps, "robots": len(models),
                             "dtype": args.precision, "batched_speedup": value / pv}
        log("plain path %.3e voxel-steps/s (%.2f s) -> batched speedup %.1fx" % (pv, pdt, value / pv))
    if args.serving_seconds > 0:
        srv, reqs = serving_run(prec, args.serving_seconds, args.serving_rate)
        # the plain path on the same requests: each alone, one step per launch sequence
        sub = [m for _, m, _ in reqs[:8]]
        pv, _ = plain_path(sub, prec, 20)
        srv["plain_path"] = {"value": pv, "unit": "voxel-steps/s", "requests": len(sub), "steps": 20}
        srv["speedup_vs_plain"] = srv["value"] / pv
        out["serving"] = srv
        log("serving: %s" % srv)
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import ref_available
        kind = "ref" if ref_available() else "oracle"
        threads = min(16, os.cpu_count() or 1)
        sample = models[: args.cpu_robots]
        cv, cdt, cfit = cpu_baseline(sample, args.cpu_steps, threads, kind)
        out["cpu_baseline"] = {"value": cv, "unit": "voxel-steps/s", "cores": threads,
                               "kind": "reference" if kind == "ref" else "port",
                               "sample": "%d whole robots (the batch's first) x %d steps, FP64 CPU engine, one robot "
                                         "per thread" % (len(sample), args.cpu_steps), "seconds": cdt}
        log("cpu baseline %.3e voxel-steps/s on %d threads (%.1f s, %s)" % (cv, threads, cdt, kind))
        par = {"robots": len(sample), "steps": args.cpu_steps, "tolerance": 0.005}
        for name, p in (("f64", _abi.PREC_F64), ("f32", _abi.PREC_F32)):
            g = gpu_fitness(sample, p, args.cpu_steps)
            rel = [abs(a - c) / abs(c) if c else abs(a) for a, c in zip(g, cfit)]
            par[name] = {"max_rel_fitness_err": max(rel), "within": int(sum(r <= 0.005 for r in rel))}
        out["fitness_parity"] = par
        log("fitness parity vs CPU engine: %s" % par)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
