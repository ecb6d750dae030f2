"""Benchmark of the PSTD update loop on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload lshape] [--no-cpu]

A step is one time step of the scene: 6 RK stages (fused line-derivative +
update launches), PML attenuation and receiver sampling, openPSTD's
SingleThreadSolver::compute_timestep (kernel/Solver.cpp:71-93).  The default
workload is the L-shaped three-domain scene (3 x 256x128 cells + PML, seed 0,
2000 steps).  Prints one JSON line.

value = cell-RK-stage updates/s = (cells of all non-rigid domains, PML
included) * 6 * steps / wall time of the timed steps, state resident on the
GPU.  roofline: algorithmic HBM bytes of one stage launch (48 B per cell:
read p0, vx, vy of the previous stage and vx, vy, px, py at t_n, write vx,
vy, px, py, p0) over the GPU time per stage launch, measured with a HIP event
pair on the solver stream around the timed steps.  roofline.traffic: measured
HBM bytes per stage launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by the
same counters' reading for a copy of known size), read from the committed
profile summary of the workload when there is one.
cpu_baseline: the CPU oracle (FP32 restatement of the reference solver with
OpenMP over calc tasks) on this host, over a bounded sample of the same scene.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-RK-stage updates/sec (GEMM+batched vs plain FFT) at ≤1e-4 rel-L2 vs FP64"
UNIT = "cell-RK-stage updates/s"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_STAGE = 48


def workload(name, steps):
    from openpstd_amd import workloads as W
    if name == "lshape":
        return W.lshape(0, steps), ("L-shaped 3-domain scene, 3 x 256x128 cells + 50-cell PML, edge absorption "
                                "from {0,0.5,1} (seed 0), 4 receivers, FP32 FFT path")
    if name == "room":
        return W.room(steps), "single 100x50-cell domain + 50-cell PML, centre source, 1 receiver, FP32 FFT path"
    if name == "square4096":
        return W.square(0, steps), "single 4096x4096-cell domain + 50-cell PML, 16 receivers, FP32 FFT path"
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(conf, cells, seconds=10.0, max_steps=400):
    """FP32 oracle on the host cores: MultiThreadSolver-style OpenMP over the
    calc tasks, timed over a bounded number of steps of the same scene."""
    from oracle import binding as ob
    threads = min(16, os.cpu_count() or 1)
    o = ob.OracleScene(conf, 32, threads)
    o.step(2)  # warm-up: FFT plans, wavenumber tables
    done, t0 = 0, time.perf_counter()
    while done < max_steps:
        o.step(5)
        done += 5
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": cells * 6 * done / dt, "unit": UNIT, "cores": threads, "kind": "port",
            "sample": f"{done} time steps of the same scene after 2 warm-up steps ({dt:.1f} s), "
                      f"FP32 oracle, {threads} OpenMP threads"}


def measured_traffic(name):
    """HBM bytes per stage launch from profiles/stage_kernel_<workload>_traffic.md
    (line 'traffic_bytes_per_launch: N'), or None."""
    path = os.path.join(ROOT, "profiles", f"stage_kernel_{name}_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def copy_calibration(mib):
    """A plain elementwise copy (torch.mul by 1, one read and one write of
    `mib` MiB) for scaling the HBM byte counters."""
    import torch
    a = torch.rand(mib * (1 << 18), device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": a.numel() * 4, "bytes_written": b.numel() * 4,
                                           "launches": 5}}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="lshape", choices=["lshape", "room", "square4096"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a copy of MIB MiB (for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the PSTD loop here runs on a single MI355X; "
              "sharding across GPUs is not built (one process, one device).", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    import openpstd_amd as P

    total = args.warmup + args.steps
    conf, desc = workload(args.workload, total)
    kernel = P.PSTDKernel(GPU=True)
    kernel.initialize_kernel(conf)
    dev = kernel.prepare()
    cells = dev.cells_per_stage

    dev.advance_async(args.warmup)
    dev.synchronize()
    t0 = time.perf_counter()
    # HIP events on the solver stream around the timed steps; returns after they finished
    mean_ms, launches = dev.profile_stages(args.steps)
    t1 = time.perf_counter()
    wall = t1 - t0
    value = cells * 6 * args.steps / wall
    achieved = BYTES_PER_CELL_STAGE * cells / (mean_ms * 1e-3) / 1e9
    traffic = measured_traffic(args.workload)

    infos = [kernel.domain_info(i) for i in range(kernel.domain_count())]
    user = sum(d["nx"] * d["ny"] for d in infos if not d["is_pml"])
    out = {
        "metric": METRIC,
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc, "cells_per_stage": cells, "user_cells": user, "domains": len(infos),
                   "rk_stages_per_step": 6, "path": "FP32 FFT (fused LDS line derivative + RK update)"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "stage kernel (one launch per RK stage)", "mean_launch_us": mean_ms * 1e3,
                     "launches_timed": launches, "bytes_per_launch": BYTES_PER_CELL_STAGE * cells},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(conf, cells)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
