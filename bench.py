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
--plain-compare N (batches): the first N scenes of the batch also run one at a
time, each on its own solver in the FP32 FFT form (the plain path), and
"plain_sequential" reports their cell-RK-stage/s (GPU time of the timed stage
launches only, host gaps between scenes excluded) and batched / plain.
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
MFMA_F16_PEAK_TFLOPS = 2500.0  # dense FP16 MFMA
BYTES_PER_CELL_STAGE = 48


WORKLOADS = ["lshape", "room", "square4096", "batch64", "sources256"]
DEFAULT_STEPS = {"lshape": 2000, "room": 500, "square4096": 100, "batch64": 100, "sources256": 10}
DEFAULT_WARMUP = {"lshape": 50, "room": 50, "square4096": 10, "batch64": 10, "sources256": 3}


def workload(name, steps):
    """(list of configurations, description); several configurations are
    advanced together as one batch."""
    from openpstd_amd import workloads as W
    if name == "lshape":
        return [W.lshape(0, steps)], ("L-shaped 3-domain scene, 3 x 256x128 cells + 50-cell PML, edge absorption "
                                      "from {0,0.5,1} (seed 0), 4 receivers")
    if name == "room":
        return [W.room(steps)], "single 100x50-cell domain + 50-cell PML, centre source, 1 receiver"
    if name == "square4096":
        return [W.square(0, steps, receiver_radius=400)], ("single 4096x4096-cell domain + 50-cell PML, centre "
                                                           "source, 16 receivers within 400 cells of it")
    if name == "sources256":
        return W.source_batch(0, steps), ("one 1024x1024-cell geometry (two 512x1024 domains + 50-cell PML) with 256 "
                                          "seeded source positions advanced together, 8 receivers each")
    if name == "batch64":
        return W.scene_batch(0, steps), ("64 independent scenes (seed 0) of 1-4 rectangular domains, sides 64-1024 "
                                         "cells, edge absorption U[0,1], 1 source + 1 receiver each, advanced "
                                         "together (one launch per RK stage)")
    raise SystemExit(f"unknown workload {name}")


def plain_sequential(confs, n, steps):
    """The first n scenes advanced one at a time (one solver each, FP32 FFT
    form, autotuned like the batch): summed cell-stages over summed GPU time of
    their timed stage launches."""
    import openpstd_amd as P
    gpu_s, cell_stages = 0.0, 0
    for conf in confs[:n]:
        k = P.PSTDKernel(GPU=True)
        k.initialize_kernel(conf)
        k.set_form("fft")
        dev = k.prepare()
        dev.advance_async(2)
        dev.synchronize()
        ms, launches = dev.profile_stages(steps)
        gpu_s += ms * 1e-3 * launches
        cell_stages += dev.cells_per_stage * launches
        del dev, k
    return {"value": cell_stages / gpu_s, "unit": UNIT, "scenes": min(n, len(confs)), "steps_per_scene": steps,
            "form": "fft"}


def _oracle_rate(confs, threads, seconds, max_steps):
    """cell-RK-stage/s of the FP32 oracle with `threads` OpenMP threads over a
    bounded sample: whole steps of one scene (or 2 steps of the first scenes
    of a batch) until `seconds` of timed work."""
    from oracle import binding as ob
    cell_steps, elapsed, steps_run, scenes = 0, 0.0, 0, 0
    for conf in confs:
        o = ob.OracleScene(conf, 32, threads)
        infos = [o.domain_info(i) for i in range(o.n_domains)]
        cells = sum(d["nx"] * d["ny"] for d in infos if not d["is_rigid"])
        o.step(1)  # warm-up: FFT tables
        per = 5 if len(confs) == 1 else 2
        done, t0 = 0, time.perf_counter()
        while done < max_steps:
            o.step(per)
            done += per
            if len(confs) > 1 or time.perf_counter() - t0 >= seconds:
                break
        elapsed += time.perf_counter() - t0
        cell_steps += cells * done
        steps_run += done
        scenes += 1
        if elapsed >= seconds:
            break
    what = (f"{steps_run} time steps of the same scene after 1 warm-up step" if len(confs) == 1 else
            f"2 time steps (after 1 warm-up step) of each of the first {scenes} of {len(confs)} scenes")
    return cell_steps * 6 / elapsed, f"{what} ({elapsed:.1f} s timed), FP32 oracle, {threads} OpenMP thread(s)"


def cpu_baseline(confs, seconds=10.0, max_steps=400):
    """FP32 oracle on the host cores: MultiThreadSolver-style OpenMP over the
    calc tasks (the reported value), and single-threaded as
    SingleThreadSolver, each timed over a bounded sample of the same
    workload."""
    threads = min(16, os.cpu_count() or 1)
    value, sample = _oracle_rate(confs, threads, seconds, max_steps)
    single, single_sample = _oracle_rate(confs, 1, seconds / 2, max_steps)
    return {"value": value, "unit": UNIT, "cores": threads, "kind": "port", "sample": sample,
            "single_thread": {"value": single, "cores": 1, "sample": single_sample}}


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
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default depends on the workload)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default depends on the workload)")
    ap.add_argument("--workload", default="lshape", choices=WORKLOADS)
    ap.add_argument("--graph-steps", type=int, default=0, help="replay blocks of this many steps as a hipGraph")
    ap.add_argument("--form", default="fft", choices=["fft", "gemm"],
                    help="line derivative form: FP32 FFT, or dense matrices on MFMA (FP16 hi+lo split)")
    ap.add_argument("--tile-order", default="auto", choices=["auto", "xcd", "plain"],
                    help="workgroup-to-tile order (default: the autotuner's choice)")
    ap.add_argument("--flavour", default="auto", choices=["auto", "early", "late", "mid"],
                    help="stage-kernel flavour (default: the autotuner's choice)")
    ap.add_argument("--tile-threads", default="auto", choices=["auto", "128", "256", "512"],
                    help="block size for lines of FFT length <= 2048 (default: the autotuner's choice)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-compare", type=int, default=0, metavar="N",
                    help="batches: also time the first N scenes one at a time in the plain FFT form")
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

    steps = args.steps if args.steps is not None else DEFAULT_STEPS[args.workload]
    warmup = args.warmup if args.warmup is not None else DEFAULT_WARMUP[args.workload]
    confs, desc = workload(args.workload, warmup + steps)
    if len(confs) == 1:
        kernel = P.PSTDKernel(GPU=True)
        kernel.initialize_kernel(confs[0])
        kernel.set_form(args.form)
        kernels = [kernel]
        dev = kernel.prepare()
    else:
        dev = P.BatchSolver(confs, form=args.form)
        kernels = dev.kernels
    dev.set_graph_steps(args.graph_steps)
    if args.tile_order != "auto":
        dev.set_tile_order(args.tile_order == "xcd")
    if args.flavour != "auto":
        dev.set_kernel_flavour(args.flavour)
    if args.tile_threads != "auto":
        dev.set_tile_threads(int(args.tile_threads))
    tuning = dev.tuning()
    cells = dev.cells_per_stage

    dev.advance_async(warmup)
    dev.synchronize()
    t0 = time.perf_counter()
    # HIP events on the solver stream around the timed steps; returns after they finished
    mean_ms, launches = dev.profile_stages(steps)
    t1 = time.perf_counter()
    wall = t1 - t0
    value = cells * 6 * steps / wall
    achieved = BYTES_PER_CELL_STAGE * cells / (mean_ms * 1e-3) / 1e9
    traffic = measured_traffic(args.workload)

    if args.form == "fft":
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                    "kernel": "stage kernel (one launch per RK stage)", "mean_launch_us": mean_ms * 1e3,
                    "launches_timed": launches, "bytes_per_launch": BYTES_PER_CELL_STAGE * cells}
    else:
        flops, outs = dev.gemm_work()
        tflops = flops / (mean_ms * 1e-3) / 1e12
        roofline = {"bound": "mfma", "achieved": tflops, "peak": MFMA_F16_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": tflops / MFMA_F16_PEAK_TFLOPS, "traffic": None,
                    "kernel": "GEMM-form stage launches (MFMA flops of the GEMM segments over the whole stage time)",
                    "mean_launch_us": mean_ms * 1e3, "launches_timed": launches, "flops_per_stage": flops,
                    "gemm_line_outputs_per_stage": outs}
    infos = [k.domain_info(i) for k in kernels for i in range(k.domain_count())]
    user = sum(d["nx"] * d["ny"] for d in infos if not d["is_pml"])
    out = {
        "metric": METRIC,
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.form == "fft" else "f16x2-split/f32-accumulate",
        "data": "synthetic",
        "config": {"workload": desc, "scenes": len(confs), "cells_per_stage": cells, "user_cells": user,
                   "domains": len(infos), "rk_stages_per_step": 6, "graph_steps": args.graph_steps,
                   "launch_shape": {"flavour": ["early", "late", "mid"][tuning["lean"]], "xcd_order": tuning["xcd_order"],
                                    "tile_threads": tuning["tile_threads"],
                                    "autotuned": args.tile_order == "auto" and args.flavour == "auto"
                                    and args.tile_threads == "auto"},
                   "path": ("FP32 FFT (fused LDS line derivative + RK update)" if args.form == "fft" else
                            "dense derivative matrices on MFMA, FP16 hi+lo split operands, FP32 accumulate "
                            "(short segments in the FP32 FFT form)")},
        "roofline": roofline,
    }
    if args.plain_compare and len(confs) > 1:
        plain = plain_sequential(confs, args.plain_compare, min(steps, 20))
        plain["batched_over_plain"] = value / plain["value"]
        out["plain_sequential"] = plain
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(confs)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
