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

Unless --block is given, the preferred workgroup size is the autotuner's choice
(vx3h_autotune) on this same workload, reported under `autotune` with the whole
trial table; resident launches run 1000 steps unless --steps-per-launch is given.

Printed: one JSON line with voxel-steps/s (value), the roofline of the dominant
kernel (HIP events on the engine's stream; for the resident kernel also its FP64
rate against FP64 peak from the committed counter profile), the plain
one-simulation-per-launch HIP path on the same robots, the other configurations as
single simulations (6^3 cube, 64x64x32 slab with self-collision, 2,048 arena balls
in FP64 and FP32), the serving stream (capacity, then Poisson arrivals at 50% and
80% of it with p50/p99 latency, and the plain path on the same requests), the CPU
engine timed on this host's cores on a bounded sample of whole robots, and the
final fitness of every robot of the batch against the CPU engine.
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
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 (vector and matrix)


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


def measured_per_launch(fname, dtype, spl):
    """A per-launch figure measured with rocprofv3 and kept in profiles/<fname> as lines
    "<dtype> <steps per launch> <value>": the least-squares line a + b*steps through the
    measured launch lengths, evaluated at spl (exact when spl was measured)."""
    import numpy as np
    f = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(f):
        return None
    pts = []
    for line in open(f):
        w = line.split()
        if len(w) == 3 and w[0] == dtype and w[1].isdigit():
            pts.append((int(w[1]), float(w[2])))
    if not pts:
        return None
    for n, v in pts:
        if n == spl:
            return v
    if len(set(n for n, _ in pts)) == 1:
        return pts[0][1] if len(pts) == 1 else float(np.mean([v for _, v in pts]))
    x = np.array([n for n, _ in pts], float)
    y = np.array([v for _, v in pts], float)
    b_, a_ = np.polyfit(x, y, 1)
    return float(a_ + b_ * spl)


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


def serve_stream(precision, reqs, quantum=500):
    """Runs an arrival stream [(arrival time, model, horizon)] through the
    continuous-batching server in real time; latency is submit to result."""
    import numpy as np
    from gpuvoxels_amd import engine
    sv = engine.Server(precision=precision, quantum_steps=quantum)
    lat, vsteps, i = [], 0.0, 0
    t0 = sv.clock()
    first = last = None
    while True:
        now = sv.clock() - t0
        while i < len(reqs) and reqs[i][0] <= now:
            sv.submit(reqs[i][1], reqs[i][2])
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
    return {"requests": len(reqs), "value": vsteps / span, "unit": "voxel-steps/s", "seconds": span,
            "voxels_min_max": [int(min(m.n_voxels for _, m, _ in reqs)), int(max(m.n_voxels for _, m, _ in reqs))],
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99))}


def serving_legs(precision, seconds, burst, seed=4):
    """The serving stream (robots of log-uniform 100..50,000 voxels, horizons uniform in
    2,000..20,000 steps).  Capacity: `burst` requests submitted at once (the server
    saturated).  Then Poisson arrivals for `seconds` at 50% and 80% of that capacity
    (rate = fraction x capacity / mean voxel-steps per request), each reporting
    throughput and p50/p99 latency, and the plain one-simulation-per-launch path on the
    burst's requests (measured first, 100 steps)."""
    import numpy as np
    from gpuvoxels_amd import workloads
    rng = np.random.default_rng(seed + 1000)
    reqs = [(0.0,) + workloads.serving_request(rng, name="burst%04d" % i) for i in range(burst)]
    work = float(np.mean([m.n_voxels * h for _, m, h in reqs]))
    # the plain path first: it is bound by host launch overhead, and measured after the
    # serving legs in the same process it has read up to half this rate
    pv, _ = plain_path([m for _, m, _ in reqs], precision, 100)
    cap = serve_stream(precision, reqs)
    out = {"dtype": "f64" if precision == 1 else "f32", "capacity": cap, "mean_voxel_steps_per_request": work}
    log("serving capacity: %.3e voxel-steps/s (%d requests at once)" % (cap["value"], burst))
    for frac in (0.5, 0.8):
        rate = frac * cap["value"] / work
        leg = serve_stream(precision, workloads.serving_stream(seconds, rate, seed=seed))
        leg.update({"load_fraction": frac, "rate_per_s": rate, "arrival_seconds": seconds})
        out["load_%d" % int(frac * 100)] = leg
        log("serving at %d%% load (%.1f req/s): %s" % (frac * 100, rate, leg))
    out["plain_path"] = {"value": pv, "unit": "voxel-steps/s", "requests": len(reqs), "steps": 100}
    out["capacity_vs_plain"] = cap["value"] / pv
    return out


def config_leg(model, precision, steps, interval=0, kernel_steps=50, warmup=10):
    """One configuration (a model, or a list of models batched together): admitted,
    `warmup` steps, then `steps` timed steps bracketed by stream synchronization;
    per-pass kernel durations from HIP events over `kernel_steps` further steps, with
    the link and voxel passes priced at their algorithmic bytes (FP64: 196 B/link,
    232 B/voxel; FP32: 124 B each)."""
    from gpuvoxels_amd import _abi, engine
    f64 = precision == _abi.PREC_F64
    models = model if isinstance(model, list) else [model]
    b = engine.Batch(precision=precision, path=_abi.PATH_AUTO, collision_rebuild_interval=interval)
    t0 = time.perf_counter()
    slots = [b.admit(m) for m in models]
    s = slots[0]
    admit_s = time.perf_counter() - t0
    V = sum(m.n_voxels for m in models)
    L = int(sum(len(engine.topology(m)[0]) for m in models))
    b.step(warmup)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    dt = time.perf_counter() - t0
    sts = b.poll()
    st = sts[s]
    wg = [b.workgroups(x) for x in slots]
    kt = b.kernel_times(kernel_steps) if kernel_steps else {}
    b.close()
    out = {"value": V * steps / dt, "unit": "voxel-steps/s", "voxels": V, "links": L, "steps": steps,
           "sims": len(models), "ms_per_step": dt / steps * 1e3, "dtype": "f64" if f64 else "f32", "admit_s": admit_s,
           "state": int(st.state), "steps_done": int(min(sts[x].steps for x in slots)), "kernel_ms": kt,
           "resident_workgroups": int(sum(wg)), "tiled_sims": int(sum(w > 1 for w in wg)),
           "alg_frac_per_step": bytes_per_step(V, L, f64) / (dt / steps) / 1e9 / HBM_PEAK_GBS}
    if "links" in kt:
        out["links_hbm_frac"] = (196 if f64 else 124) * L / (kt["links"] * 1e-3) / 1e9 / HBM_PEAK_GBS
        out["links_hbm_frac_at_124B"] = 124 * L / (kt["links"] * 1e-3) / 1e9 / HBM_PEAK_GBS
    if "voxels" in kt:
        out["voxels_hbm_frac"] = (232 if f64 else 124) * V / (kt["voxels"] * 1e-3) / 1e9 / HBM_PEAK_GBS
    return out


def config_legs(args):
    """The configuration legs named in args.leg_names, each a batch of its own."""
    import numpy as np
    from gpuvoxels_amd import _abi, workloads
    legs = {}
    names = set(args.leg_names.split(","))
    if "C0" in names:
        m0 = workloads.cube(6, steps=0)
        legs["C0_single_cube_f64"] = config_leg(m0, _abi.PREC_F64, args.leg_steps or 10000, kernel_steps=0)
    if "C2" in names:
        m2 = workloads.slab(64, 64, 32, steps=0)
        legs["C2_slab_f64"] = config_leg(m2, _abi.PREC_F64, args.leg_steps or 5000)
    if "C1c" in names:
        rc = workloads.robots(args.robots, seed=args.seed, steps=0)
        for r in rc:   # the Voxelyze3 default: self-collision on
            r.collisions_enabled = 1
        legs["C1_robots_with_collisions_f64"] = config_leg(rc, _abi.PREC_F64, args.leg_steps or 2000, kernel_steps=100)
    if "C3" in names or "C3f32" in names:
        m3 = workloads.arena_balls(steps=0)
        if "C3" in names:
            legs["C3_arena_balls_f64"] = config_leg(m3, _abi.PREC_F64, args.leg_steps or 3000, interval=8)
        if "C3f32" in names:
            legs["C3_arena_balls_f32"] = config_leg(m3, _abi.PREC_F32, args.leg_steps or 3000, interval=8)
    if "C4t" in names or "C4tf32" in names:
        # the serving stream's large requests: 15 robots on 30^3 grids (about 13k voxels
        # each), too large for one workgroup, on the tiled resident kernel
        rng = np.random.default_rng(args.seed + 30)
        mt = [workloads.robot(rng, grid=30, steps=0, name="big%02d" % i) for i in range(15)]
        if "C4t" in names:
            legs["C4_large_robots_tiled_f64"] = config_leg(mt, _abi.PREC_F64, args.leg_steps or 4000, kernel_steps=1000)
        if "C4tf32" in names:
            legs["C4_large_robots_tiled_f32"] = config_leg(mt, _abi.PREC_F32, args.leg_steps or 4000, kernel_steps=1000)
    for k, v in legs.items():
        log("leg %s: %.3e voxel-steps/s, %.4f ms/step, kernels %s" % (k, v["value"], v["ms_per_step"], v["kernel_ms"]))
    return legs


def precision_horizon(models, horizons, cpu_robots, threads, kind):
    """Fitness error against horizon on the robot batch, as evidence for the precision
    choice.  At each horizon (steps), relative fitness differences of:
      f32_vs_f64     FP32 state vs FP64 state, every robot;
      f64_vs_cpu     FP64 state vs the CPU engine, the first cpu_robots robots;
      f64_perturbed_1e-9 / 1e-14  FP64 runs whose voxels start with random velocities
                     of that many m/s (N(0, eps) per axis) vs the unperturbed FP64 run:
                     how far the workload itself carries a perturbation of about FP32
                     (1e-9 m/s against velocities of ~1e-2 m/s) or FP64 rounding size."""
    import copy
    import numpy as np
    from gpuvoxels_amd import _abi, engine

    def gpu_curve(ms, precision):
        b = engine.Batch(precision=precision, path=_abi.PATH_AUTO)
        slots = [b.admit(m) for m in ms]
        out, done = [], 0
        for h in horizons:
            b.step(h - done)
            done = h
            st = b.poll()
            out.append(np.array([st[s].fitness for s in slots]))
        b.close()
        return out

    def stats(a, ref):
        rel = np.abs(a - ref) / np.abs(ref)
        return {"max": float(rel.max()), "median": float(np.median(rel)), "within_0.5%": int((rel <= 0.005).sum()),
                "robots": int(len(rel))}

    f64 = gpu_curve(models, _abi.PREC_F64)
    f32 = gpu_curve(models, _abi.PREC_F32)
    res = {"horizons": horizons, "robots": len(models)}
    res["f32_vs_f64"] = [stats(a, r) for a, r in zip(f32, f64)]
    for eps in (1e-9, 1e-14):
        rng = np.random.default_rng(123)
        pert = []
        for m in models:
            q = copy.copy(m)
            q.init_velocity = rng.normal(0.0, eps, size=(m.n_voxels, 3))
            pert.append(q)
        res["f64_perturbed_%g" % eps] = [stats(a, r) for a, r in zip(gpu_curve(pert, _abi.PREC_F64), f64)]
    # the CPU engine at the same horizons
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from concurrent.futures import ThreadPoolExecutor
    from oracle_lib import OSim, load
    load(kind)

    def cpu_one(m):
        o = OSim(m, kind=kind)
        fits, done = [], 0
        for h in horizons:
            o.step(h - done)
            done = h
            fits.append(o.status().fitness)
        return fits
    with ThreadPoolExecutor(max_workers=threads) as ex:
        cf = np.array(list(ex.map(cpu_one, models[:cpu_robots])))
    res["f64_vs_cpu"] = [stats(f64[i][:cpu_robots], cf[:, i]) for i in range(len(horizons))]
    res["f32_vs_cpu"] = [stats(f32[i][:cpu_robots], cf[:, i]) for i in range(len(horizons))]
    res["cpu_kind"] = "reference" if kind == "ref" else "port"
    return res


def host_threads():
    """Worker threads for the CPU engine: OMP_NUM_THREADS when set (the GPU host's
    share of cores), else the cores this process may run on."""
    n = os.environ.get("OMP_NUM_THREADS")
    if n and n.isdigit() and int(n) > 0:
        return int(n)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


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
                    help="steps of one resident kernel launch; default 1000")
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
    ap.add_argument("--serving-seconds", type=float, default=8.0,
                    help="arrival window of each serving load level (0 = skip serving)")
    ap.add_argument("--serving-burst", type=int, default=100,
                    help="requests submitted at once to measure the serving capacity (also the plain path's requests)")
    ap.add_argument("--fitness-robots", type=int, default=512,
                    help="robots whose final fitness is checked against the CPU engine (0 = skip)")
    ap.add_argument("--precision-horizon", action="store_true",
                    help="instead of the benchmark: fitness error against horizon (FP32 vs FP64, FP64 vs the CPU "
                         "engine, FP64 under tiny initial-velocity perturbations) on the robot batch")
    ap.add_argument("--no-legs", dest="legs", action="store_false",
                    help="skip the single-cube, slab and arena-balls configuration legs")
    ap.add_argument("--legs", dest="leg_names", default="C0,C2,C1c,C3,C3f32,C4t,C4tf32",
                    help="comma-separated configuration legs to run: C0 (single cube), C2 (slab), C1c (the robots "
                         "with self-collision), C3 (arena balls, FP64), C3f32 (arena balls, FP32), C4t / C4tf32 (15 "
                         "robots of about 13k voxels on the tiled resident kernel, FP64 / FP32)")
    ap.add_argument("--legs-only", action="store_true",
                    help="run only the configuration legs named by --legs (one JSON object of legs, no headline): "
                         "one batch per leg, for profiling a leg in a process of its own")
    ap.add_argument("--serving-only", action="store_true",
                    help="run only the serving legs (one JSON object), for profiling them in a process of their own")
    ap.add_argument("--leg-steps", type=int, default=0,
                    help="timed steps of every configuration leg (0 = each leg's own: 10000, 5000, 2000, 3000)")
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
    if args.legs_only:
        legs = config_legs(args)
        print(json.dumps({"config_legs": legs}))
        return
    if args.serving_only:
        prec = _abi.PREC_F64 if args.precision == "f64" else _abi.PREC_F32
        print(json.dumps({"serving": serving_legs(prec, args.serving_seconds, args.serving_burst)}))
        return
    models = workloads.robots(args.robots, seed=args.seed, steps=0)
    if args.precision_horizon:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import ref_available
        kind = "ref" if ref_available() else "oracle"
        print(json.dumps(precision_horizon(models, [1000, 2000, 5000, 10000, 20000], args.cpu_robots, host_threads(),
                                           kind)))
        return
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
        # launches of 1000 steps unless given: the tuner's launch lengths differ by a few
        # percent at most, and a fixed length keeps the launch durations comparable with
        # the committed kernel profiles (profiles/kernel_trace_resident_512robots.txt)
        spl_ = args.steps_per_launch if args.steps_per_launch is not None else 1000
        rec["used"] = {"block_size": block, "steps_per_launch": spl_}
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
        traffic = measured_per_launch("hbm_traffic_resident_512robots.txt", args.precision, spl)
    flops_step = measured_per_launch("fp64_flops_resident_512robots.txt", args.precision, 1) if dom == "resident" else None
    per_step_survey = 124 * (V + L)
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
            "kernel": dom, "alg_bytes_per_launch": ab, "steps_per_launch": spl, "avg_launch_ms": kt[dom],
            "alg_bytes_per_step": per_step, "kernel_ms": kt,
            # the same launches priced at 124 B per voxel and per link (FP32 structure-of-arrays bytes)
            "frac_at_124B": (per_step_survey * spl / (kt[dom] * 1e-3) / 1e9 / HBM_PEAK_GBS) if dom == "resident" else None}
    if dom == "resident":
        # the resident kernel keeps the state on chip: its measured HBM traffic per launch
        # is ~0.1% of the algorithmic bytes, so what bounds it is on chip (FP64 issue, LDS
        # and barrier waits; counters in profiles/), priced here against FP64 peak
        roof["limiter"] = "on-chip: FP64 VALU issue, LDS access and barrier waits (HBM traffic per launch is measured, not bounding)"
        if flops_step:
            tf = flops_step * spl / (kt[dom] * 1e-3) / 1e12
            roof["compute"] = {"kind": "fp64 vector", "flops_per_step": flops_step, "achieved": tf,
                               "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tf / FP64_PEAK_TFLOPS}

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
    if args.legs:
        out["config_legs"] = config_legs(args)
    if args.serving_seconds > 0:
        out["serving"] = serving_legs(prec, args.serving_seconds, args.serving_burst)
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import ref_available
        kind = "ref" if ref_available() else "oracle"
        threads = host_threads()
        sample = models[: args.cpu_robots]
        cv, cdt, _ = cpu_baseline(sample, args.cpu_steps, threads, kind)
        out["cpu_baseline"] = {"value": cv, "unit": "voxel-steps/s", "cores": threads,
                               "kind": "reference" if kind == "ref" else "port",
                               "sample": "%d whole robots (the batch's first) x %d steps, FP64 CPU engine, one robot "
                                         "per thread" % (len(sample), args.cpu_steps), "seconds": cdt,
                               "host_cpus": os.cpu_count()}
        log("cpu baseline %.3e voxel-steps/s on %d threads (%.1f s, %s)" % (cv, threads, cdt, kind))
        if args.fitness_robots:
            # every robot of the batch over the full horizon: CPU engine vs both GPU precisions
            chk = models[: args.fitness_robots]
            _, fdt, cfit = cpu_baseline(chk, args.cpu_steps, threads, kind)
            par = {"robots": len(chk), "steps": args.cpu_steps, "tolerance": 0.005, "cpu_seconds": fdt}
            for name, p in (("f64", _abi.PREC_F64), ("f32", _abi.PREC_F32)):
                g = gpu_fitness(chk, p, args.cpu_steps)
                rel = [abs(a - c) / abs(c) if c else abs(a) for a, c in zip(g, cfit)]
                par[name] = {"max_rel_fitness_err": max(rel), "median_rel_fitness_err": sorted(rel)[len(rel) // 2],
                             "within": int(sum(r <= 0.005 for r in rel))}
            out["fitness_parity"] = par
            log("fitness parity vs CPU engine: %s" % par)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
