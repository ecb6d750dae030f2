"""Benchmark of the PSTD update loop on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload sources256] [--form gemm_f16]

A step is one time step of the scene(s): 6 RK stages (fused line-derivative +
update launches), PML attenuation and receiver sampling, openPSTD's
SingleThreadSolver::compute_timestep (kernel/Solver.cpp:71-93).

Default workload: one 1024 x 1024-cell geometry (two 512 x 1024 domains +
50-cell PML) with 256 seeded source positions advanced together, 8 receivers
each, in the GEMM form (dense derivative matrices on MFMA, FP16 hi+lo split
operands, FP32 accumulation; the thin PML strips in the FFT form) -- the
"GEMM+batched" side of the metric.  Prints one JSON line.

value = cell-RK-stage updates/s = (cells of all non-rigid domains, PML
included) * 6 * steps / wall time of the timed steps, state resident on the
GPU.  The timed steps are issued eagerly with a HIP event pair on the solver
stream around every launch of each kernel class (pstd_ctx_profile_kernels).
roofline (GEMM form): algorithmic MFMA flops of one GEMM-kernel launch (3
split products x 2 x K per line output) over that kernel's mean launch time
from those events; peak 2.5 PFLOP/s dense FP16/BF16.  (FFT form: algorithmic
HBM bytes of one stage launch, 48 B per cell, over the stage kernel's mean
launch time; peak 8 TB/s.)  roofline.traffic: HBM bytes per launch of that
kernel from the committed rocprofv3 counter summary of the workload (FETCH_SIZE
+ WRITE_SIZE scaled by the same counters' reading for a copy of known size),
or null.
plain_fft: the first N scenes advanced one at a time, each on its own solver in
the plain FP32 FFT form (wall time of each scene's step loop, synchronised on
both sides); vs_plain = value / that rate.
accuracy: the first sources of the workload (as many as the FP64 fixture
holds) advanced for the full 3000 steps in
the benchmarked form against the committed FP64 oracle series
(tests/golden/sources1024_receivers.npz): receiver rel-L2 and max-abs per
source.
secondary: the L-shaped three-domain scene in the FP32 FFT form (HBM roofline).
cpu_baseline: the CPU oracle (FP32 restatement of the reference solver with
OpenMP over calc tasks, openPSTD's MultiThreadSolver structure) on every core
this process may use, over a bounded sample of the same workload.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-RK-stage updates/sec (GEMM+batched vs plain FFT) at ≤1e-4 rel-L2 vs FP64"
UNIT = "cell-RK-stage updates/s"
HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = 2500.0  # dense FP16 / BF16 MFMA
BYTES_PER_CELL_STAGE = 48

WORKLOADS = ["sources256", "lshape", "room", "square4096", "batch64"]
DEFAULT_STEPS = {"lshape": 2000, "room": 500, "square4096": 100, "batch64": 100, "sources256": 20}
DEFAULT_WARMUP = {"lshape": 50, "room": 50, "square4096": 10, "batch64": 10, "sources256": 5}
FORM_DTYPE = {"fft": "f32", "gemm_f16": "f16x2-split/f32-accumulate", "gemm_bf16": "bf16x2-split/f32-accumulate"}


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


def make_solver(confs, form):
    import openpstd_amd as P
    if len(confs) == 1:
        kernel = P.PSTDKernel(GPU=True)
        kernel.initialize_kernel(confs[0])
        kernel.set_form(form)
        return kernel.prepare(), [kernel]
    dev = P.BatchSolver(confs, form=form)
    return dev, dev.kernels


def plain_fft(confs, n, steps, warmup=2):
    """The first n scenes advanced one at a time (one solver each, FP32 FFT
    form, autotuned like the batch): summed cell-stages over the summed wall
    time of their step loops (and over the GPU time of their stage launches)."""
    import openpstd_amd as P
    wall, gpu_s, cell_stages = 0.0, 0.0, 0
    for conf in confs[:n]:
        k = P.PSTDKernel(GPU=True)
        k.initialize_kernel(conf)
        k.set_form("fft")
        dev = k.prepare()
        dev.advance_async(warmup + steps)  # also captures the solver's hipGraph, outside the timed loop
        dev.synchronize()
        t0 = time.perf_counter()
        dev.advance_async(steps)
        dev.synchronize()
        wall += time.perf_counter() - t0
        ms, launches = dev.profile_stages(steps)
        gpu_s += ms * 1e-3 * launches
        cell_stages += dev.cells_per_stage * 6 * steps
        del dev, k
    return {"value": cell_stages / wall, "unit": UNIT, "value_gpu_time": cell_stages / gpu_s,
            "scenes": min(n, len(confs)), "steps_per_scene": steps, "form": "fft",
            "method": "each scene on its own solver, FP32 FFT form; cell-stages over the wall time of its "
                      "step loop (advance + synchronise); value_gpu_time over the GPU time of its stage launches"}


def accuracy(form, steps=3000):
    """The first sources of the 256-source workload (as many as the FP64 fixture
    holds), full 3000 steps, against the committed oracle series: per source
    rel-L2 and max-abs."""
    import numpy as np
    import openpstd_amd as P
    from openpstd_amd import workloads as W
    path = os.path.join(ROOT, "tests", "golden", "sources1024_receivers.npz")
    if not os.path.exists(path):
        return None
    g = np.load(path)
    n_src = len([k for k in g.files if k.startswith("fp64_source")])
    confs = W.source_batch(0, steps, n_sources=n_src)
    b = P.BatchSolver(confs, form=form)
    got = b.advance(steps)
    out = {}
    for i in range(n_src):
        s, want = b.scene_samples(got, i), g[f"fp64_source{i}"]
        out[f"source{i}"] = {"rel_l2": float(np.linalg.norm(s - want) / np.linalg.norm(want)),
                             "max_abs": float(np.abs(s - want).max())}
    return {"reference": "FP64 oracle, tests/golden/sources1024_receivers.npz", "steps": steps, "form": form,
            "per_source": out, "max_rel_l2": max(v["rel_l2"] for v in out.values())}


def measured_traffic(name, kernel):
    """HBM bytes per launch of `kernel` from profiles/<kernel>_<workload>_traffic.md
    (line 'traffic_bytes_per_launch: N'), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{name}_traffic.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1]), os.path.relpath(path, ROOT)
    return None, None


def secondary_lshape(steps=400, warmup=50):
    """The L-shaped scene in the FP32 FFT form: cell-stage/s and HBM roofline."""
    dev, _ = make_solver(workload("lshape", 2 * steps + warmup)[0], "fft")
    cells = dev.cells_per_stage
    phase(f"warmup {warmup} steps, then {steps} timed steps")
    dev.advance_async(warmup)
    dev.synchronize()
    dev.capture_graph()
    t0 = time.perf_counter()
    dev.advance_async(steps)
    dev.synchronize()
    wall = time.perf_counter() - t0
    mean_ms, launches = dev.profile_stages(steps)
    achieved = BYTES_PER_CELL_STAGE * cells / (mean_ms * 1e-3) / 1e9
    return {"workload": workload("lshape", 1)[1], "form": "fft", "value": cells * 6 * steps / wall, "unit": UNIT,
            "steps": steps, "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                         "frac": achieved / HBM_PEAK_GBS, "mean_launch_us": mean_ms * 1e3,
                                         "note": "about 3 workgroups per CU: launch- and latency-limited"}}


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


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(confs, seconds=10.0, max_steps=400):
    """FP32 oracle on all host cores available to this process:
    MultiThreadSolver-style OpenMP over the calc tasks (the reported value), and
    single-threaded as SingleThreadSolver, each timed over a bounded sample of
    the same workload."""
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count() or 1
    value, sample = _oracle_rate(confs, threads, seconds, max_steps)
    single, single_sample = _oracle_rate(confs, 1, seconds / 2, max_steps)
    return {"value": value, "unit": UNIT, "cores": threads, "kind": "port", "sample": sample,
            "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "fft": "the oracle's own radix-2 FFT (FFTW is not available)",
            "single_thread": {"value": single, "cores": 1, "sample": single_sample}}


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


def phase(msg):
    """One progress line on stderr per bench phase (the JSON result alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default depends on the workload)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default depends on the workload)")
    ap.add_argument("--workload", default="sources256", choices=WORKLOADS)
    ap.add_argument("--form", default="gemm_f16", choices=["fft", "gemm_f16", "gemm_bf16", "auto"],
                    help="line derivative form: FP32 FFT, or dense matrices on MFMA (FP16 / BF16 hi+lo split)")
    ap.add_argument("--graph-steps", type=int, default=None,
                    help="replay blocks of this many steps as a hipGraph (default 20; kernel-class timing runs "
                         "eagerly)")
    ap.add_argument("--tile-order", default="auto", choices=["auto", "xcd", "plain"],
                    help="workgroup-to-tile order (default: the autotuner's choice)")
    ap.add_argument("--flavour", default="auto", choices=["auto", "early", "late", "mid"],
                    help="FFT stage-kernel flavour (default: the autotuner's choice)")
    ap.add_argument("--tile-threads", default="auto", choices=["auto", "128", "256", "512"],
                    help="block size for lines of FFT length <= 2048 (default: the autotuner's choice)")
    ap.add_argument("--receivers", default="physical", choices=["physical", "spectral"],
                    help="receiver reading: the receiver's cell, or spectral interpolation at its location")
    ap.add_argument("--overlap", action="store_true",
                    help="GEMM form: run a stage's FFT-form tiles on a second stream beside the GEMM kernel")
    ap.add_argument("--plain-compare", type=int, default=8, metavar="N",
                    help="batches: time the first N scenes one at a time in the plain FFT form (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
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

    steps = args.steps if args.steps is not None else DEFAULT_STEPS[args.workload]
    warmup = args.warmup if args.warmup is not None else DEFAULT_WARMUP[args.workload]
    confs, desc = workload(args.workload, warmup + 5 * steps + 4)
    if args.receivers == "spectral":
        from openpstd_amd import abi_types as T
        for c in confs:
            c.receiver_mode = T.PSTD_RECEIVER_SPECTRAL
        desc += "; spectral-interpolation receivers"
        args.no_accuracy = True  # the FP64 fixture holds nearest-cell series
    phase(f"{args.workload}: {len(confs)} scenes, form {args.form}: building plans and autotuning")
    dev, kernels = make_solver(confs, args.form)
    form = dev.form()  # with --form auto: the form the upload found faster
    dev.set_graph_steps(20 if args.graph_steps is None else args.graph_steps)
    if args.overlap:
        dev.set_stage_overlap(True)
    if args.tile_order != "auto":
        dev.set_tile_order(args.tile_order == "xcd")
    if args.flavour != "auto":
        dev.set_kernel_flavour(args.flavour)
    if args.tile_threads != "auto":
        dev.set_tile_threads(int(args.tile_threads))
    tuning = dev.tuning()
    cells = dev.cells_per_stage

    phase(f"warmup {warmup} steps, then {steps} timed steps")
    dev.advance_async(warmup)
    dev.synchronize()
    if warmup > 0:
        dev.capture_graph()  # the step graph exists before the timed loop
    t0 = time.perf_counter()
    dev.advance_async(steps)
    dev.synchronize()
    wall = time.perf_counter() - t0
    value = cells * 6 * steps / wall
    # the same number of steps again, issued eagerly with a HIP event pair
    # around every launch: per-kernel GPU times for the roofline
    prof = dev.profile_kernels(steps)
    stage_ms = prof["total"][0] / prof["total"][1]

    if form == "fft":
        cls = max(("fft_small", "fft_4096", "fft_8192"), key=lambda c: prof[c][0])
        mean_ms = prof[cls][0] / prof[cls][1]
        kname = {"fft_small": "stage_small_kernel", "fft_4096": "stage_big_kernel_4096",
                 "fft_8192": "stage_split_kernel"}[cls]
        # one event pair around as many steps again, replayed like the timed
        # ones (per-launch pairs inflate launches of a few microseconds): the
        # stage time of the HBM roofline
        stage_ms, _ = dev.profile_stages(steps)
        achieved = BYTES_PER_CELL_STAGE * cells / (stage_ms * 1e-3) / 1e9
        traffic, tsrc = measured_traffic(args.workload, kname)
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": tsrc,
                    "kernel": f"the stage's FFT-form launches (dominant: {kname}, mean event-timed launch "
                              f"{mean_ms * 1e3:.1f} us); 48 B per cell of every non-rigid domain over the stage time",
                    "mean_launch_us": mean_ms * 1e3, "stage_us": stage_ms * 1e3,
                    "bytes_per_stage": BYTES_PER_CELL_STAGE * cells,
                    "limiter": "latency and issue, not HBM bandwidth: one line pair per tile, its gather, "
                               "transform passes and update in sequence (counters in profiles/)"}
    else:
        flops, outs = dev.gemm_work()
        g_ms, g_n = prof["gemm"]
        mean_ms = g_ms / g_n
        tflops = flops / (mean_ms * 1e-3) / 1e12
        traffic, tsrc = measured_traffic(args.workload, "stage_gemm_kernel")
        roofline = {"bound": "mfma", "achieved": tflops, "peak": MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": tflops / MFMA_PEAK_TFLOPS, "traffic": traffic, "traffic_source": tsrc,
                    "kernel": "stage_gemm_kernel (one launch per RK stage: every GEMM-form tile of every scene)",
                    "mean_launch_us": mean_ms * 1e3, "launches_timed": g_n, "flops_per_launch": flops,
                    "gemm_line_outputs_per_launch": outs,
                    "limiter": "latency and issue: the kernel's MFMA loop alone (operands from LDS) runs at 86 % of "
                               "the peak; load latency, the FP32->16-bit split with its LDS staging and the epilogue "
                               "make up the rest (profiles/stage_gemm_kernel_sources256_counters.md)",
                    "whole_stage": {"stage_us": stage_ms * 1e3,
                                    "achieved_tflops": flops / (stage_ms * 1e-3) / 1e12,
                                    "frac": flops / (stage_ms * 1e-3) / 1e12 / MFMA_PEAK_TFLOPS}}
    kernel_times = {c: {"ms_total": v[0], "launches": v[1]} for c, v in prof.items() if v[1]}
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
        "dtype": FORM_DTYPE[form],
        "data": "synthetic",
        "config": {"workload": desc, "scenes": len(confs), "cells_per_stage": cells, "user_cells": user,
                   "domains": len(infos), "rk_stages_per_step": 6, "form": form,
                   "form_requested": args.form,
                   "launch_shape": {"flavour": ["early", "late", "mid"][tuning["lean"]],
                                    "xcd_order": tuning["xcd_order"], "tile_threads": tuning["tile_threads"],
                                    "autotuned": args.tile_order == "auto" and args.flavour == "auto"
                                    and args.tile_threads == "auto",
                                    "autotune_trial_steps": tuning["trial_steps"],
                                    "autotune_cached": tuning["cached"],
                                    "graph_steps": 20 if args.graph_steps is None else args.graph_steps},
                   "path": ("FP32 FFT (fused LDS line derivative + RK update)" if form == "fft" else
                            "dense derivative matrices on MFMA, 16-bit hi+lo split operands staged from FP32 "
                            "in the kernel, FP32 accumulate, fused RK update (thin PML strips in the FP32 FFT "
                            "form)")},
        "roofline": roofline,
        "kernel_times": kernel_times,
        "timing": ("value: wall time of the timed steps, replayed as hipGraphs of graph_steps steps captured "
                   "before the loop; kernel_times (and the GEMM roofline): the same number of steps again, issued "
                   "eagerly with a HIP event pair on the solver stream around every launch; FFT roofline: one "
                   "event pair around as many steps again, replayed as the timed ones"),
    }
    del dev
    phase(f"timed: {wall / steps * 1e3:.2f} ms per step")
    if args.plain_compare and len(confs) > 1:
        phase(f"plain FFT form, {args.plain_compare} scenes one at a time")
        plain = plain_fft(confs, args.plain_compare, steps)
        out["plain_fft"] = plain
        out["vs_plain"] = value / plain["value"]
    if args.workload == "sources256" and not args.no_accuracy:
        phase("accuracy: 3000 steps vs the FP64 fixture")
        out["accuracy"] = accuracy(form)
    if not args.no_secondary and args.workload != "lshape":
        phase("secondary: L-shape, FFT form")
        out["secondary"] = secondary_lshape()
    if not args.no_cpu:
        phase("CPU baseline (oracle, all cores of this process)")
        out["cpu_baseline"] = cpu_baseline(confs)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
