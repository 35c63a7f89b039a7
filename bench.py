"""Headline benchmark: D3Q19 BGK+Guo stream+collide, FP32, 512x128x128.

Workload (the MLUPS / %HBM-roofline metric's configuration): periodic x,
bounce-back walls at y in {0,127} and z in {0,127}, Guo body force 1e-6 along
x, tau = 0.6, from rest (f = w_i).  One step = one fused AA stream+collide
launch over all 8,388,608 cells.  Inputs are resident in HBM before the
timed region; the timed region is bracketed by stream synchronisation, and
the device time comes from HIP events recorded on the lattice's own stream.

Algorithmic bytes per cell and step: 2 * 19 * 4 = 152 (one read and one
write of every population).  Prints one JSON line.

--gpus N > 1 exits: this engine runs one lattice on one GPU.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0


def build_c2(lattice, storage, ext, g=1e-6, tau=0.6):
    import numpy as np
    from lbm_mirror_amd import abi
    from lbm_mirror_amd.lattice import Lattice
    from helpers import duct_flags_3d, c2_ops
    lat = Lattice(lattice, storage, ext)
    lat.set_flags(duct_flags_3d(*ext))
    lat.set_periodic(1)
    for fl, op in c2_ops(tau, g).items():
        lat.set_operator(fl, op)
    lat.init_equilibrium(1.0, (0.0, 0.0, 0.0))
    return lat


def cpu_baseline(ext, steps=2, threads=None):
    """The oracle (the reference's arithmetic restated in C++, OpenMP over
    cells) on the host cores, on a bounded sample of the same workload."""
    import numpy as np
    from lbm_mirror_amd import abi
    from oracle import oracle as orc
    from helpers import duct_flags_3d, c2_ops
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    s = orc.Sim(abi.D3Q19, abi.F32, ext)
    s.set_flags(duct_flags_3d(*ext))
    s.set_periodic(1)
    for fl, op in c2_ops(0.6, 1e-6).items():
        s.set_operator(fl, op)
    s.init_equilibrium(1.0, (0.0, 0.0, 0.0))
    s.step(1, threads=threads)  # warm-up (first-touch)
    t0 = time.perf_counter()
    s.step(steps, threads=threads)
    dt = time.perf_counter() - t0
    cells = ext[0] * ext[1] * ext[2]
    return {"value": round(cells * steps / dt / 1e6, 2), "unit": "MLUPS", "cores": threads,
            "kind": "port",
            "sample": f"D3Q19 FP32 {ext[0]}x{ext[1]}x{ext[2]}, {steps} steps after 1 warm-up, "
                      f"oracle (reference arithmetic, OpenMP) on {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--lattice", default="D3Q19", choices=["D3Q19", "D3Q27"])
    ap.add_argument("--storage", default="F32", choices=["F32", "F64", "F16S", "BF16S"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs one lattice on one GPU; multi-GPU "
                          "sharding is a design note only (x-slab halo exchange)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import lbm_mirror_amd as lm
    from lbm_mirror_amd import abi
    L = lm.load_library()
    if L.lbm_device_count() < 1:
        raise SystemExit("no GPU visible to liblbm_mirror_amd.so")
    lat_id = {"D3Q19": abi.D3Q19, "D3Q27": abi.D3Q27}[args.lattice]
    st_id = {"F32": abi.F32, "F64": abi.F64, "F16S": abi.F16S, "BF16S": abi.BF16S}[args.storage]
    ext = (512, 128, 128) if args.lattice == "D3Q19" else (256, 256, 256)
    Q = 19 if args.lattice == "D3Q19" else 27
    sbytes = {"F32": 4, "F64": 8, "F16S": 2, "BF16S": 2}[args.storage]
    bytes_per_cell = 2 * Q * sbytes

    print(f"[bench] building {args.lattice} {args.storage} {ext}", file=sys.stderr, flush=True)
    lat = build_c2(lat_id, st_id, ext)
    cells = ext[0] * ext[1] * ext[2]
    if args.warmup:
        lat.step(args.warmup, t0=0)
    lat.sync()
    print(f"[bench] warm-up done; timing {args.steps} steps", file=sys.stderr, flush=True)
    w0 = time.perf_counter()
    lat.step(args.steps, t0=args.warmup)
    lat.sync()
    wall = time.perf_counter() - w0
    dev_ms = lat.last_step_ms()
    ms_per_step = dev_ms / args.steps
    mlups = cells / (ms_per_step * 1e-3) / 1e6
    achieved = cells * bytes_per_cell / (ms_per_step * 1e-3) / 1e9
    rho, u = lat.get_macros()
    import numpy as np
    if not np.all(np.isfinite(rho)):
        raise SystemExit("non-finite density after the timed run")
    traffic = None
    tr_path = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tr_path):
        tr = json.load(open(tr_path))
        key = f"{args.lattice}_{args.storage}"
        traffic = tr.get(key)
    out = {
        "metric": "MLUPS (D3Q19 FP32 stream+collide)" if args.lattice == "D3Q19" else
                  "MLUPS (D3Q27 FP32 stream+collide)",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"F32": "f32", "F64": "f64", "F16S": "f16-storage/f32", "BF16S": "bf16-storage/f32"}[args.storage],
        "data": "synthetic",
        "config": {"workload": f"{args.lattice} BGK+Guo periodic-x channel, BB walls y/z, "
                               f"{ext[0]}x{ext[1]}x{ext[2]}, tau 0.6, g 1e-6",
                   "lattice": list(ext), "bytes_per_cell_step": bytes_per_cell,
                   "wall_s": round(wall, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
    }
    lat.close()
    if not args.no_cpu:
        print("[bench] CPU baseline (oracle)", file=sys.stderr, flush=True)
        try:
            out["cpu_baseline"] = cpu_baseline(ext, steps=args.cpu_steps)
        except Exception as e:  # the baseline is reported, never the product
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
