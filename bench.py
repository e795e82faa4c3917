"""Benchmark of the MI355X dycore large step (plain FP64).

Workload: stratified flow over a bell-shaped hill, 433 x 433 x 50 mass points,
dx = 3 km, stretched eta, dt = 18 s, 6 acoustic substeps (the configuration
the headline cell-substeps/s metric is quoted on).  Synthetic seeded initial
condition (wrf_gpu_amd.ic.stratified_terrain_3d); inputs are resident in HBM
before the timed region.

A "step" is one RK3 large step (3 stages, 1 + 3 + 6 acoustic substeps).
Throughput counts namelist substeps: cell-substeps/s = N * ns * steps / wall.

Prints ONE JSON line with the metric, the dominant kernel's roofline (its
algorithmic bytes per launch divided by its average launch duration, timed
with HIP events on the launch stream over the timed region) and a CPU
baseline (the oracle on a bounded sample of the same workload, this host's
cores).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "cell-substeps/s, chosen form vs plain FP64 at stated err; % HBM roofline"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per mass cell for one launch of each kernel: every
# 3-D array the kernel must read or write once (FP64), neighbour re-reads and
# scratch excluded.
ALG_ARRAYS_PER_CELL = {
    # reads u_work, v_work, u, v, th_work, t_tend, theta, ww, rom, w_work,
    # rw_tend, ph_tend, ph_1, phb, ph_work, c2a, alt, a, alpha, gamma, t_2ave,
    # w_save, pm1 (23); writes th_work, ww, w_work, ph_work, t_2ave, pm1, p, al (8)
    "k_column": 31,
    # advance_uv for u and v: reads u_work, v_work, u_tend, v_tend, p, ph_work,
    # pb, al, alt, php; writes u_work, v_work
    "k_uv": 12,
}
# Per-large-step compulsory traffic model (cells x 8 B x passes):
# 3 stages x 55 stage passes + S(ns) x 45 substep passes, S(ns) = 1 + ns//2 + ns.


def step_bytes_per_cell(ns: int) -> int:
    s = 1 + max(1, ns // 2) + ns
    return 8 * (3 * 55 + s * 45)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=433)
    ap.add_argument("--ny", type=int, default=433)
    ap.add_argument("--nz", type=int, default=50)
    ap.add_argument("--cpu-baseline-steps", type=int, default=1,
                    help="large steps of the oracle timed on the host (0 skips)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all cores in this process's affinity")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (counter calibration)")
    return ap.parse_args()


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU; the dycore shards as described in DESIGN.md "
                                   "but multi-GPU runs are not part of it", "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import Dycore, state_to_device

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)

    t0 = time.time()
    case = ic.stratified_terrain_3d(nx=args.nx, ny=args.ny, nz=args.nz)
    m, nl = case.metrics, case.namelist
    n_cells = m.n_cells
    print(f"[bench] IC {m.nx}x{m.ny}x{m.nz} built in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    dev = torch.device("cuda:0")
    state = state_to_device(case.state, dev)
    dyn = Dycore(m, nl)
    stream = torch.cuda.current_stream(dev)

    if args.calibration_copy_mib > 0:
        # a plain copy of known size: HBM byte counters read on it calibrate
        # the counters read on the dycore kernels
        n = args.calibration_copy_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # warmup
    dyn.step(state, args.warmup)
    torch.cuda.synchronize()
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)

    # timed region, dominant-kernel events recorded on the launch stream
    profiled = ("k_column", "k_uv")
    dyn.profile(profiled)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    dyn.step(state, args.steps)
    end.record(stream)
    torch.cuda.synchronize()
    wall_ms = start.elapsed_time(end)
    kern = dyn.profile_read()
    dyn.profile(())
    print(f"[bench] timed {args.steps} steps: {wall_ms:.1f} ms", file=sys.stderr, flush=True)

    finite = all(bool(torch.isfinite(t).all()) for t in state.values())
    ms_per_step = wall_ms / max(args.steps, 1)
    ns = nl.acoustic_substeps
    value = n_cells * ns * args.steps / (wall_ms / 1e3)
    executed = n_cells * (1 + max(1, ns // 2) + ns) * args.steps / (wall_ms / 1e3)

    # roofline of the dominant profiled kernel
    per_kernel = {}
    for name, (ms, cnt) in kern.items():
        avg_ms = ms / cnt
        alg = ALG_ARRAYS_PER_CELL[name] * 8 * n_cells
        per_kernel[name] = {"avg_ms": avg_ms, "launches": cnt, "share_of_wall": ms / wall_ms,
                            "alg_bytes": alg, "achieved_GBs": alg / (avg_ms * 1e-3) / 1e9}
    dom = max(per_kernel, key=lambda k: per_kernel[k]["avg_ms"] * per_kernel[k]["launches"])
    d = per_kernel[dom]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(d["achieved_GBs"], 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(d["achieved_GBs"] / HBM_PEAK_GBS, 4), "traffic": None,
                "alg_bytes_per_launch": d["alg_bytes"], "avg_launch_ms": round(d["avg_ms"], 4)}
    step_bytes = step_bytes_per_cell(ns) * n_cells
    step_roofline = {"model_bytes_per_step": step_bytes,
                     "achieved": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1), "unit": "GB/s",
                     "frac": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    # CPU baseline: the oracle (a C restatement of the reference path) on a
    # bounded sample of the same workload
    cpu = None
    if args.cpu_baseline_steps > 0:
        from oracle import oracle

        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        oracle.build()
        t1 = time.time()
        oracle.run(m, nl, case.state, args.cpu_baseline_steps, threads=threads)
        cpu_s = time.time() - t1
        cpu = {"value": n_cells * ns * args.cpu_baseline_steps / cpu_s, "unit": "cell-substeps/s",
               "cores": threads, "kind": "port",
               "sample": f"{args.cpu_baseline_steps} large step(s) of the same {m.nx}x{m.ny}x{m.nz} workload "
                         f"from the initial state, CPU oracle (C, OpenMP), {cpu_s:.1f} s"}
        print(f"[bench] cpu baseline {cpu_s:.1f}s on {threads} threads", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain FP64 stratified flow over terrain {m.nx}x{m.ny}x{m.nz}, dx=3 km, "
                               f"stretched eta, dt={nl.dt_s:g} s, ns={ns}",
                   "n_cells": n_cells, "form": "plain_fp64", "parallelism": "single GPU"},
        "executed_cell_substeps_per_s": executed,
        "roofline": roofline,
        "step_roofline": step_roofline,
        "kernels": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in per_kernel.items()},
        "cpu_baseline": cpu,
        "finite": finite,
    }
    print(json.dumps(out))
    dyn.close()


if __name__ == "__main__":
    main()
