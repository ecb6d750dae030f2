"""Writes the committed fixtures of tests/golden/.

1. Known-answer data held by openPSTD's own unit tests, read as text from the
   reference checkout (only where it is present; the fixtures are committed):
     window_w32_pe70.json      test/Kernel/kernel_functions.cpp:147
     kat_reference.json        next_2_power, get_rho_array, wavenumbers, Gaussian
2. Long-run oracle series (FP64 CPU oracle), too slow to recompute in a test:
     room_receivers.npz, lshape_receivers.npz, square1024_receivers.npz
     (a 1024 x 1024-cell domain whose lines fill the GEMM form's blocks, 300
     steps, 16 receivers within 120 cells of the source),
     sources1024_receivers.npz (the 256-source geometry: 1024 x 1024 cells as
     two 512 x 1024 domains, its first two seeded source positions with 8
     seeded receivers each, the full 3000 steps, FP64),
     lshape_frames.npz (the L-shaped scene's p0 of its three user domains after
     the full 2000 steps, FP64 and FP32),
     square4096_receivers.npz (the 4096 x 4096-cell domain of the bench, 16
     seeded receivers within 200 cells of the centre source, 500 steps, FP64),
     batch64_receivers.npz (four scenes of the 64-scene batch, seed 0: 19 with
     one domain, 56 with two, 40 with three, 50 with four; 2000 steps, FP64),
     square4096_long_receivers.npz (the 4096 x 4096-cell domain at its
     configured 1000 steps, 16 receivers within 400 cells of the source, FP64;
     written every 100 steps),
     lshape_spectral_receivers.npz (the L-shape with ten spectral-interpolation
     receivers, 500 steps, FP64 and FP32)

    python tests/golden/make_fixtures.py [--kat] [--series] [--square] [--sources] [--frames] [--square4096]
                                         [--batch64] [--square4096-long] [--lshape-spectral] [--sources-more N]
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, ROOT)


def kat():
    src = open(os.path.join(REF, "test/Kernel/kernel_functions.cpp")).read()
    m = re.search(r"window_verify << ([^;]*);", src)
    vals = [float(v) for v in m.group(1).split(",")]
    json.dump({"source": "test/Kernel/kernel_functions.cpp:147 (window_verify, w=32, patch error 70)",
               "values": vals}, open(os.path.join(HERE, "window_w32_pe70.json"), "w"), indent=1)
    wc = open(os.path.join(REF, "test/Kernel/WisdomCache.cpp")).read()
    checks = re.findall(r"is_approx\(discr1\.(\w+)(?:\.(real|imag)\(\))?\.coeff\((\d+)\), (-?[\d.]+)\)", wc)
    data = {
        "next_2_power": {"source": "test/Kernel/kernel_functions.cpp:47-51",
                         "cases": [[42, 64], [3.4, 4], [0.1, 1]]},
        "rho_array": {"source": "test/Kernel/kernel_functions.cpp:53-74",
                      "cases": [{"rho": [1e10, 1.2, 1.2], "pressure": [1, -1, 0, 0, 2, 0, 1, 1],
                                 "velocity": [-1, 1, 0, 0, 0, 2, 1, 1]},
                                {"rho": [1.2, 1.2, 1.2], "pressure": [0, 0, 0, 0, 1, 1, 1, 1],
                                 "velocity": [0, 0, 0, 0, 1, 1, 1, 1]}]},
        "discretization": {"source": "test/Kernel/WisdomCache.cpp:48-98", "dx": 0.2, "n": 115,
                           "table_len": 128, "larger": [[178, 256], [227, 256]], "max_k": 15.8,
                           "checks": [[name, part or "", int(idx), float(v)] for name, part, idx, v in checks]},
        "gaussian": {"source": "test/Kernel/Domain.cpp:47-70, 199-214", "center_rc": [25, 20],
                     "neighbour_value": 0.706947},
        "scene_counts": {"source": "test/Kernel/Scene.cpp:75-123, test/Kernel/Domain.cpp:97-106,168-173",
                         "reflecting_alpha01": {"domains": 9, "pml": 8, "secondary": 4, "pml_side": 50},
                         "domain_test": {"domains": 9, "size_x": 250, "bottom_right_y": 300,
                                         "neighbours_no_pml": 0, "neighbours_all": 4}},
    }
    json.dump(data, open(os.path.join(HERE, "kat_reference.json"), "w"), indent=1)
    print("wrote KAT fixtures")


def series(which=("room", "lshape")):
    from oracle import binding as ob
    from openpstd_amd import workloads as W
    cases = {"room": (W.room(500), 500), "lshape": (W.lshape(0, 2000), 2000),
             "square1024": (W.square(0, 300, n=1024, n_receivers=16, receiver_radius=120), 300)}
    for name in which:
        conf, steps = cases[name]
        out = {}
        for prec in (64, 32):
            o = ob.OracleScene(conf, prec, threads=os.cpu_count() or 1)
            out[f"fp{prec}"] = o.step(steps)
            print(name, prec, "done", flush=True)
        np.savez(os.path.join(HERE, f"{name}_receivers.npz"), **out)


def lshape_spectral(steps=500):
    """FP64 and FP32 oracle series of the L-shape with spectral-interpolation receivers."""
    from oracle import binding as ob
    from openpstd_amd import workloads as W
    conf = W.lshape_spectral(steps)
    out = {}
    for prec in (64, 32):
        out[f"fp{prec}"] = ob.OracleScene(conf, prec, threads=os.cpu_count() or 1).step(steps)
        print("lshape_spectral", prec, "done", flush=True)
    np.savez(os.path.join(HERE, "lshape_spectral_receivers.npz"), **out)


def sources(n_sources=2, steps=3000):
    from oracle import binding as ob
    from openpstd_amd import workloads as W
    threads = int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1))
    out = {}
    for i, conf in enumerate(W.source_batch(0, steps, n_sources=n_sources)):
        out[f"fp64_source{i}"] = ob.OracleScene(conf, 64, threads=threads).step(steps)
        print("source", i, "done", flush=True)
    np.savez(os.path.join(HERE, "sources1024_receivers.npz"), **out)


def sources_more(n_total, steps=3000):
    """Adds FP64 series of the next seeded source positions of the same
    geometry (sources 0 .. n_total-1 of one seed: the first ones are those
    already stored) to sources1024_receivers.npz, saving after each (about
    50 CPU minutes per source on 8 cores)."""
    from oracle import binding as ob
    from openpstd_amd import workloads as W
    threads = int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1))
    path = os.path.join(HERE, "sources1024_receivers.npz")
    out = dict(np.load(path)) if os.path.exists(path) else {}
    for i, conf in enumerate(W.source_batch(0, steps, n_sources=n_total)):
        if f"fp64_source{i}" in out:
            continue
        out[f"fp64_source{i}"] = ob.OracleScene(conf, 64, threads=threads).step(steps)
        np.savez(path, **out)
        print("source", i, "done", flush=True)


def square4096(steps=500):
    from oracle import binding as ob
    from openpstd_amd import workloads as W
    threads = int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1))
    conf = W.square(0, steps, n=4096, n_receivers=16, receiver_radius=200)
    np.savez(os.path.join(HERE, "square4096_receivers.npz"), fp64=ob.OracleScene(conf, 64, threads=threads).step(steps))
    print("square4096 done", flush=True)


def square4096_long(steps=1000, chunk=100):
    """The 4096 x 4096-cell domain at its configured 1000 steps, 16 seeded
    receivers within 400 cells of the centre source (FP64).  About 18 s of
    CPU per step on 8 cores: the file is rewritten every `chunk` steps with
    the steps done so far."""
    from oracle import binding as ob
    from openpstd_amd import workloads as W
    threads = int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1))
    conf = W.square(0, steps, n=4096, n_receivers=16, receiver_radius=400)
    o = ob.OracleScene(conf, 64, threads=threads)
    parts = []
    for done in range(0, steps, chunk):
        parts.append(o.step(min(chunk, steps - done)))
        np.savez(os.path.join(HERE, "square4096_long_receivers.npz"), fp64=np.concatenate(parts))
        print("square4096_long", done + len(parts[-1]), "steps", flush=True)


BATCH64_SCENES = (19, 56, 40, 50)


def batch64(steps=2000):
    from oracle import binding as ob
    from openpstd_amd import workloads as W
    threads = int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1))
    confs = W.scene_batch(0, steps)
    out = {}
    for i in BATCH64_SCENES:
        out[f"fp64_scene{i}"] = ob.OracleScene(confs[i], 64, threads=threads).step(steps)
        print("scene", i, "done", flush=True)
    np.savez(os.path.join(HERE, "batch64_receivers.npz"), **out)


def frames(steps=2000):
    from oracle import binding as ob
    from openpstd_amd import workloads as W
    threads = int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1))
    out = {}
    for prec in (64, 32):
        o = ob.OracleScene(W.lshape(0, steps), prec, threads=threads)
        o.step(steps)
        for i in range(o.n_domains):
            if not o.domain_info(i)["is_pml"]:
                out[f"fp{prec}_d{i}"] = o.field(i, "p0").astype(np.float64 if prec == 64 else np.float32)
        print("frames", prec, "done", flush=True)
    np.savez(os.path.join(HERE, "lshape_frames.npz"), **out)


if __name__ == "__main__":
    args = sys.argv[1:] or ["--kat", "--series"]
    if "--kat" in args:
        kat()
    if "--series" in args:
        series()
    if "--square" in args:
        series(("square1024",))
    if "--sources" in args:
        sources()
    if "--frames" in args:
        frames()
    if "--square4096" in args:
        square4096()
    if "--batch64" in args:
        batch64()
    if "--square4096-long" in args:
        square4096_long()
    if "--lshape-spectral" in args:
        lshape_spectral()
    if "--sources-more" in args:
        sources_more(int(args[args.index("--sources-more") + 1]))
