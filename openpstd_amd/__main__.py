"""Command line of openpstd_amd, after openPSTD's CLI (CLI/main.cpp).

    python -m openpstd_amd create scene.pstd          # a new file with the default scene
    python -m openpstd_amd run -g scene.pstd          # run the scene on the GPU, results into the file
    python -m openpstd_amd info scene.pstd            # scene and result counts
    python -m openpstd_amd export scene.pstd -F npz   # results as NumPy arrays (HDF5 layout)

`run` follows RunCommand (CLI/main.cpp:365-464): -g selects the GPU solver
(this library's); without it openPSTD's CPU solvers would run, which this
library does not provide (an error).  --receivers picks how receivers read
the pressure (physical: the receiver's cell, reference: openPSTD's
transposed read, spectral: spectral interpolation at the exact location).
`export` follows ExportCommand (CLI/main.cpp:476-600) with the layout of
openPSTD's HDF5 export (shared/export/HDF5Export.cpp:20-80): arrays named
frame/<domain>/<frame> (dimensions (nx, ny) over the frame's row-major data,
as openPSTD declares them) and receiver/<receiver>, in one NumPy .npz file
(<output-dir>/<output-name>.npz): the image has no HDF5 library.
"""
import argparse
import sys

from . import abi_types as T
from .kernel import KernelCallbackLog
from .pstd_file import PSTDFile

MODES = {"physical": T.PSTD_RECEIVER_PHYSICAL, "reference": T.PSTD_RECEIVER_REFERENCE,
         "spectral": T.PSTD_RECEIVER_SPECTRAL}


class ConsoleLog(KernelCallbackLog):
    """CLIOutput's log half (CLI/output.cpp:41-67)."""

    def __init__(self, debug=False):
        self.debug = debug

    def Fatal(self, message): print("FATAL:", message, file=sys.stderr)
    def Error(self, message): print("ERROR:", message, file=sys.stderr)
    def Warning(self, message): print("WARNING:", message)
    def Info(self, message): print(message)

    def Debug(self, message):
        if self.debug:
            print(message)


def export(f, a):
    """HDF5Export::ExportData's datasets (shared/export/HDF5Export.cpp:20-80)."""
    import os

    import numpy as np

    from .kernel import PSTDKernel
    conf = f.GetResultsSceneConf()
    k = PSTDKernel(GPU=True)
    k.initialize_kernel(conf)  # the scene's domain sizes (get_metadata; no device work)
    meta = k.get_metadata().DomainMetadata
    out = {}
    domains = a.domains if a.domains else list(range(f.GetResultsDomainCount()))
    for d in domains:
        end = a.end_frame if a.end_frame >= 0 else f.GetResultsFrameCount(d) - 1
        for fr in range(a.start_frame, end + 1):
            out[f"frame/{d}/{fr}"] = f.GetResultsFrame(fr, d).reshape(meta[d][0], meta[d][1])
    for r in range(f.GetResultsReceiverCount()):
        out[f"receiver/{r}"] = f.GetReceiverData(r)
    os.makedirs(a.output_dir, exist_ok=True)
    np.savez(os.path.join(a.output_dir, a.output_name + ".npz"), **out)


def main(argv=None):
    p = argparse.ArgumentParser(prog="python -m openpstd_amd")
    sub = p.add_subparsers(dest="command", required=True)
    c = sub.add_parser("create", help="create a scene file with the default scene")
    c.add_argument("scene_file")
    r = sub.add_parser("run", help="run the scene of a file, results into the file")
    r.add_argument("scene_file")
    r.add_argument("-m", "--multithreaded", action="store_true", help="openPSTD's multi-threaded CPU solver")
    r.add_argument("-g", "--gpu-accelerated", action="store_true", help="the GPU solver")
    r.add_argument("--receivers", choices=sorted(MODES), default="physical")
    r.add_argument("--debug", action="store_true")
    i = sub.add_parser("info", help="scene and results of a file")
    i.add_argument("scene_file")
    e = sub.add_parser("export", help="results as arrays (openPSTD's HDF5 export layout, NumPy .npz)")
    e.add_argument("scene_file")
    e.add_argument("-F", "--format", default="npz", choices=["npz"])
    e.add_argument("-d", "--output-dir", default="./")
    e.add_argument("-n", "--output-name", default="result")
    e.add_argument("-s", "--start-frame", type=int, default=0)
    e.add_argument("-e", "--end-frame", type=int, default=-1)
    e.add_argument("-D", "--domains", type=int, nargs="*", default=None)
    a = p.parse_args(argv)
    try:
        if a.command == "create":
            PSTDFile.New(a.scene_file)
        elif a.command == "run":
            f = PSTDFile.Open(a.scene_file)
            f.run(gpu=a.gpu_accelerated, mcpu=a.multithreaded, receiver_mode=MODES[a.receivers],
                  log=ConsoleLog(a.debug))
        elif a.command == "export":
            export(PSTDFile.Open(a.scene_file), a)
        else:
            f = PSTDFile.Open(a.scene_file)
            conf, res = f.GetSceneConf(), f.GetResultsSceneConf()
            print(f"scene: {len(conf.Domains)} domains, {len(conf.Speakers)} speakers, "
                  f"{len(conf.Receivers)} receivers, {conf.num_steps() if conf.Settings.GetGridSpacing() else 0} steps")
            for d in range(len(res.Domains)):
                print(f"results: domain {d}: {f.GetResultsFrameCount(d)} frames")
            for r_ in range(len(res.Receivers)):
                print(f"results: receiver {r_}: {f.GetReceiverData(r_).size} samples")
    except Exception as e:  # the CLI prints the error and exits with 1, as RunCommand does
        print("error:", e, file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
