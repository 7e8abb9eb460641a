#!/usr/bin/env python
"""Benchmark: DM reconstruction throughput, diffraction frames/s per iteration.

Workload (the reference simulation shape): 30 x 30 raster scan, 512 x 512 detector
(lambda 0.248 nm, z 1 m, 172 um pixels => 2.82 nm object pixels, 50 nm = 17.76 px steps),
1088 x 1088 random-field object (1027 x 1027 reconstruction canvas), 110 nm Gaussian probe,
1e7 photons/frame Poisson, one probe mode.  A step is one full DM iteration over all 900
frames with inputs resident in HBM.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-frames S | --no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
KERNELS = ("row_pass", "column_pass", "inverse_row_pass", "probe_gather", "object_gather", "probe_apply",
           "object_apply", "finish")


def dm_flags(i: int) -> int:
    """IPhaser::phaseStep with the Python defaults phaseConstraint=1, updateProbe=10."""
    from ptychopy_amd import _engine as E
    f = 0
    if i <= 1:
        f |= E.FLAG_PHASE_CONSTRAINT
    if i >= 10:
        f |= E.FLAG_UPDATE_PROBE
    return f


def kernel_bytes(name: str, n: int, m: int, obj_px: int) -> float:
    """Algorithmic (compulsory) HBM bytes of one iteration's launches of each DM kernel kind."""
    px = n * m * m
    return {
        "row_pass": 16.0 * px + 8.0 * m * m + 8.0 * obj_px,     # psi in, X out; P and O once
        "column_pass": 18.0 * px,                                # X in, FP16 A in, Z out
        "inverse_row_pass": 16.0 * px,                           # Z in, psi out
        "probe_gather": 8.0 * px + 8.0 * obj_px,                 # psi in; O once
        "object_gather": 8.0 * px + 8.0 * m * m,                 # psi in; P once
        "probe_apply": 32.0 * m * m + 8.0 * m * m,
        "object_apply": 24.0 * obj_px,
        "finish": 8.0,
    }[name]


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(wl: dict, frames: int) -> dict:
    """FP64 oracle DM iteration over the first `frames` frames of the same workload."""
    import oracle
    from ptychopy_amd import _engine as E
    m = wl["m"]
    s = min(frames, wl["n"])
    amp = wl["amp_deq"][:s].astype(np.float64)
    pos = wl["pos"][:s]
    obj = np.zeros(wl["canvas"], np.complex128)
    probe = wl["probe0"].astype(np.complex128)
    psi = np.zeros((s, m, m), np.complex128)
    fac = np.array([1e-3, 1e-3])
    oracle.dm_iteration(obj, probe, psi[:2], amp[:2], pos[:2], fac)  # warm-up (thread pool, pages)
    t0 = time.perf_counter()
    oracle.dm_iteration(obj, probe, psi, amp, pos, fac, update_probe=True)
    dt = time.perf_counter() - t0
    return {"value": s / dt, "unit": "frames/s", "cores": oracle.num_threads(), "kind": "port",
            "sample": f"one FP64 oracle DM iteration (probe update on) over the first {s} of {wl['n']} "
                      f"frames, 512x512, same 1027x1027 canvas; {dt:.1f} s on {cpu_model()}"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-frames", type=int, default=384)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-iters", type=int, default=5)
    ap.add_argument("--dm-wave", type=int, default=64,
                    help="frames per DM wave (0 automatic, -1 one pass over all frames)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} / WORLD_SIZE {world} is not supported "
              "(the DM frame loop would shard by scan region with one object all-reduce per iteration; "
              "that path is not built)", file=sys.stderr)
        return 2

    import torch
    from ptychopy_amd import workloads
    from ptychopy_amd._engine import Engine

    t_gen = time.perf_counter()
    wl = workloads.dm_c1()
    t_gen = time.perf_counter() - t_gen
    n, m = wl["n"], wl["m"]
    rows, cols = wl["canvas"]
    e = Engine(n, m, (rows, cols))
    if args.dm_wave is not None:
        from ptychopy_amd._engine import TUNE_DM_WAVE
        e.set_tuning(TUNE_DM_WAVE, args.dm_wave)
    e.set_frames(wl["amp"])
    e.set_positions(wl["pos"])
    e.clear_object()
    e.init_probe(wl["beam"], wl["dx_s"])
    wl["amp_deq"] = e.get_frames()
    wl["probe0"] = e.get_probe()[0]

    it = 0
    for _ in range(args.warmup):
        it += 1
        e.dm_iteration(dm_flags(it))
    e.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        it += 1
        e.dm_iteration(dm_flags(it))
    e.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps

    # per-kernel durations from HIP events on the engine stream
    kt = np.zeros(len(KERNELS))
    for _ in range(args.kernel_iters):
        it += 1
        kt += e.dm_iteration_timed(dm_flags(it))
    kt /= args.kernel_iters
    obj_px = rows * cols
    dom = int(np.argmax(kt))
    name = KERNELS[dom]
    ach = kernel_bytes(name, n, m, obj_px) / (kt[dom] * 1e-3) / 1e9
    traffic = None
    tfile = ROOT / "profiles" / "dm_c1_traffic.json"
    if tfile.exists():
        traffic = json.loads(tfile.read_text()).get(name)
    bpf = workloads.dm_bytes_per_frame(m, 1, obj_px, n)
    it_gbs = bpf * n / (ms * 1e-3) / 1e9
    out = {
        "metric": "diffraction frames/s per iteration",
        "value": n / (ms * 1e-3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded random-field object, Gaussian probe, Poisson 1e7 photons/frame)",
        "config": {"workload": "DM 30x30 raster, 512x512 frames, 1088x1088 object (1027x1027 canvas), "
                               "1 probe mode, FP16 amplitudes, probe update on",
                   "frames": n, "frame": m, "canvas": [rows, cols], "modes": 1},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": traffic},
        "iteration_roofline": {"bytes_per_frame": bpf, "achieved": it_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": it_gbs / HBM_PEAK_GBS},
        "kernel_ms": {k: float(v) for k, v in zip(KERNELS, kt)},
        "dm_wave_frames": e.get_tuning(1),
        "setup_s": t_gen,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_frames)
    e.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
