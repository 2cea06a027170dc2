#!/usr/bin/env python3
"""Benchmark of the Voxelyze3 physics path on one MI355X.

Workload: 512 uneven robots on a 12^3 grid (Bernoulli(0.5) occupancy, largest
face-connected component, four materials, seed 0), the full physics of the
single-cube configuration (gravity, floor with static/kinetic friction, 2 Hz
phase actuation with CTE 0.01, Poisson effects), FP32 state.  One "step" is one
physics time step of every robot of the batch; the timed region holds exactly
--steps steps after --warmup untimed ones, with the batch's stream synchronized
on both sides, state resident in HBM throughout.

Printed: one JSON line with voxel-steps/s (value), the per-pass kernel roofline
of the dominant kernel (HIP events on the engine's stream), the plain
one-simulation-per-launch HIP path on the same robots, and the CPU oracle timed
on this host's cores on a bounded sample.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
B_VOXEL = 124                  # compulsory FP32 structure-of-arrays bytes per voxel per step
B_LINK = 124                   # ... per link per step


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alg_bytes(pass_name, V, L, steps_per_launch=1):
    if pass_name == "links":
        return B_LINK * L
    if pass_name == "voxels":
        return B_VOXEL * V
    if pass_name == "resident":
        return (B_VOXEL * V + B_LINK * L) * steps_per_launch
    return None


def cpu_baseline(models, steps, threads):
    """The CPU oracle (FP64 restatement of the reference CPU engine) on a bounded
    sample: `len(models)` robots x `steps` steps, one robot per thread at a time."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OSim, load
    load("oracle")
    work = list(models)
    lock = threading.Lock()
    done = [0]

    def worker():
        while True:
            with lock:
                if not work:
                    return
                m = work.pop()
            o = OSim(m)
            o.step(steps)
            with lock:
                done[0] += m.n_voxels * steps

    ts = [threading.Thread(target=worker) for _ in range(threads)]
    t0 = time.perf_counter()
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    dt = time.perf_counter() - t0
    return done[0] / dt, dt


def plain_path(models, steps):
    """One simulation per launch sequence: every robot in its own single-sim batch,
    stepped one step at a time in turn from one host thread (no batching, no graph)."""
    from gpuvoxels_amd import _abi, engine
    bs = []
    for m in models:
        b = engine.Batch(precision=_abi.PREC_F32, path=_abi.PATH_STREAM, use_graph=0)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--robots", type=int, default=512)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--path", choices=["auto", "stream", "resident"], default="auto")
    ap.add_argument("--precision", choices=["f32", "f64"], default="f32")
    ap.add_argument("--block", type=int, default=0, help="preferred resident workgroup size (0 = engine default)")
    ap.add_argument("--plain-steps", type=int, default=100)
    ap.add_argument("--cpu-robots", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=2000)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the engine shards independent simulations across GPUs with no "
              "communication, and this build runs on a single MI355X; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from gpuvoxels_amd import _abi, engine, workloads
    if engine.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    models = workloads.robots(args.robots, seed=args.seed, steps=0)
    V = sum(m.n_voxels for m in models)
    topo = [engine.topology(m) for m in models[:0]]
    del topo
    path = {"auto": _abi.PATH_AUTO, "stream": _abi.PATH_STREAM, "resident": _abi.PATH_RESIDENT}[args.path]
    prec = _abi.PREC_F64 if args.precision == "f64" else _abi.PREC_F32
    b = engine.Batch(precision=prec, path=path, use_graph=1, block_size=args.block)
    for m in models:
        b.admit(m)
    L = int(sum(len(engine.topology(m)[0]) for m in models))
    log("robots %d voxels %d links %d" % (len(models), V, L))
    b.step(args.warmup)
    b.sync()
    log("warmup done")
    t0 = time.perf_counter()
    b.step(args.steps)
    b.sync()
    dt = time.perf_counter() - t0
    value = V * args.steps / dt
    log("timed %d steps in %.3f s -> %.3e voxel-steps/s" % (args.steps, dt, value))
    st = b.poll()
    states = [s.state for s in st.values()]
    NT = 200
    kt = b.kernel_times(NT)
    log("kernel ms", kt)
    dom = max(kt, key=kt.get)
    spl = NT if dom == "resident" else 1
    ab = alg_bytes(dom, V, L, spl)
    achieved = ab / (kt[dom] * 1e-3) / 1e9 if ab else None
    roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": None,
            "alg_bytes_per_launch": ab, "steps_per_launch": spl, "avg_ms": kt[dom], "kernel_ms": kt}
    b.close()

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
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "%d uneven robots, 12^3 grid, Bernoulli(0.5) largest face-connected component, "
                               "4 materials, seed %d, floor+friction+gravity+2 Hz actuation+Poisson" % (len(models), args.seed),
                   "robots": len(models), "voxels": V, "links": L, "path": args.path,
                   "all_running": all(s == _abi.SIM_RUNNING for s in states)},
        "roofline": roof,
    }
    if not args.no_plain:
        sub = models[: min(len(models), 512)]
        pv, pdt = plain_path(sub, args.plain_steps)
        out["plain_path"] = {"value": pv, "unit": "voxel-steps/s", "steps": args.plain_steps, "robots": len(sub),
                             "batched_speedup": value / pv}
        log("plain path %.3e voxel-steps/s (%.2f s)" % (pv, pdt))
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        sample = models[: args.cpu_robots]
        cv, cdt = cpu_baseline(sample, args.cpu_steps, threads)
        out["cpu_baseline"] = {"value": cv, "unit": "voxel-steps/s", "cores": threads, "kind": "port",
                               "sample": "%d robots x %d steps of the same workload, FP64 oracle, one robot per thread"
                                         % (len(sample), args.cpu_steps), "seconds": cdt}
        log("cpu baseline %.3e voxel-steps/s on %d threads (%.1f s)" % (cv, threads, cdt))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
