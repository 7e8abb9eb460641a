#!/usr/bin/env python
"""Benchmark: reconstruction throughput, diffraction frames/s per iteration.

Default workload `mls_c2` (the configuration the headline metric is quoted on): MLs on a
51 x 51 raster (100 nm = 14 px steps), 256 x 256 frames, five incoherent Hermite-Gaussian probe
modes (power 0.6/0.2/0.1/0.06/0.04), 1e5 photons/frame, FP16 amplitudes, delta_p = 0.1; a step is
one MLs iteration over all 2601 frames (101 groups of ~26 frames) with inputs resident in HBM.
Every timed step runs the full iteration: probe and object updates, variable probe and
intensity, sub-pixel shifts and position refinement (the timed iterations are numbered from the
position-search iteration on, whatever --warmup is).

`--workload dm_c1`: DM on a 30 x 30 raster, 512 x 512 detector (lambda 0.248 nm, z 1 m, 172 um
pixels => 2.82 nm object pixels, 50 nm = 17.76 px steps), 1088 x 1088 random-field object
(1027 x 1027 reconstruction canvas), 110 nm Gaussian probe, 1e7 photons/frame, one mode; every
timed step has the probe update on and the phase constraint off.
`--workload epie_c0`: ePIE, 12 x 12 raster, 128 x 128 frames (positions/s; a dependent chain).
`--workload batch_c3`: 16 DM scans of uneven size advanced in one captured HIP graph per
iteration (and, beside it, one after another).
`--workload c4`: the learned warm start -- trains the network on simulated frames, then runs
MLs on the mls_c2 workload from a uniform and from the stitched start to a target normalised
object error (minutes; not a throughput line).

Prints one JSON line.

  python bench.py [--gpus 1] [--workload mls_c2|dm_c1|epie_c0|batch_c3|c4] [--steps K] [--warmup W] [--no-cpu]
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


def dm_flags(i: int, timed: bool = False) -> int:
    """IPhaser::phaseStep with the Python defaults phaseConstraint=1, updateProbe=10 for the
    warm-up; every timed step has the probe update on and the phase constraint off, whatever
    its iteration number."""
    from ptychopy_amd import _engine as E
    if timed:
        return E.FLAG_UPDATE_PROBE
    f = 0
    if i <= 1:
        f |= E.FLAG_PHASE_CONSTRAINT
    if i >= 10:
        f |= E.FLAG_UPDATE_PROBE
    return f


def host_info() -> dict:
    return {"nproc": os.cpu_count(), "cpu_model": cpu_model()}


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


def cpu_baseline(wl: dict, frames: int, min_seconds: float = 10.0) -> dict:
    """FP64 oracle DM iterations (probe update on) over the first `frames` frames of the same
    workload, repeated until at least min_seconds of CPU work."""
    import oracle
    m = wl["m"]
    s = min(frames, wl["n"])
    amp = wl["amp_deq"][:s].astype(np.float64)
    pos = wl["pos"][:s]
    obj = np.zeros(wl["canvas"], np.complex128)
    probe = wl["probe0"].astype(np.complex128)
    psi = np.zeros((s, m, m), np.complex128)
    fac = np.array([1e-3, 1e-3])
    oracle.dm_iteration(obj, probe, psi[:2], amp[:2], pos[:2], fac)  # warm-up (thread pool, pages)
    iters, t0 = 0, time.perf_counter()
    while True:
        oracle.dm_iteration(obj, probe, psi, amp, pos, fac, update_probe=True)
        iters += 1
        dt = time.perf_counter() - t0
        if dt >= min_seconds and iters >= 2:
            break
    return {"value": s * iters / dt, "unit": "frames/s", "cores": oracle.num_threads(), "kind": "port",
            "sample": f"{iters} FP64 oracle DM iterations (probe update on) over the first {s} of {wl['n']} "
                      f"frames, 512x512, same 1027x1027 canvas; {dt:.1f} s on {cpu_model()}", **host_info()}


def mls_cpu_baseline(W: dict, amp_deq, probe0, obj0, vmode, evo1, groups, it: int,
                     min_seconds: float = 10.0) -> dict:
    """FP64 oracle MLs iteration `it` over the first groups of the same workload: two groups time
    the rate, then a fresh state runs as many groups as take about min_seconds."""
    import oracle
    a64 = amp_deq.astype(np.float64)
    fp = (a64 ** 2).sum(axis=(1, 2))
    prm = oracle.mls_params(W["box"], iter=it)

    def run(sample):
        st = oracle.MlsState(obj0, probe0, vmode, evo1, W["pos"], a64)
        st.init(fp)
        t0 = time.perf_counter()
        st.iteration(prm, sample)
        return time.perf_counter() - t0

    dt2 = run(groups[:2])
    ng = int(min(len(groups), max(2, round(min_seconds / max(dt2 / 2, 1e-3)))))
    sample = groups[:ng]
    frames = int(sum(len(g) for g in sample))
    dt = run(sample)
    return {"value": frames / dt, "unit": "frames/s", "cores": oracle.num_threads(), "kind": "port",
            "sample": f"FP64 oracle MLs iteration {it} over its first {ng} groups ({frames} of {W['n']} frames; "
                      f"whole-canvas steps included), 256x256, 5 modes; {dt:.1f} s on {cpu_model()}",
            **host_info()}


def _base_line(args, metric, value, unit, ms, dtype, data, config) -> dict:
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": data, "config": config}


def run_epie(args) -> dict:
    """ePIE at 12 x 12 / 128^2: positions are a dependent chain, so the figure is positions/s and
    the per-position latency (no roofline fraction binds a latency-bound chain)."""
    import torch
    from ptychopy_amd import workloads
    from ptychopy_amd import _engine as E
    W = workloads.epie_c0()
    n, m = W["n"], W["m"]
    e = E.Engine(n, m, W["canvas"])
    e.set_frames(W["amp"])
    e.set_positions(W["pos"])
    e.clear_object()
    e.init_probe(W["beam"], W["dx_s"])
    amp_deq = e.get_frames()
    rng = np.random.default_rng(20261200)

    def flags(i):
        return (E.FLAG_PHASE_CONSTRAINT if i <= 1 else 0) | (E.FLAG_UPDATE_PROBE if i >= 10 else 0)

    it = 0
    for _ in range(args.warmup):
        it += 1
        e.epie_iteration(rng.permutation(n), flags(it))
    e.synchronize()
    torch.cuda.synchronize()
    orders = [rng.permutation(n) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for o in orders:  # every timed iteration updates the probe, whatever --warmup is
        it += 1
        e.epie_iteration(o, E.FLAG_UPDATE_PROBE)
    e.synchronize()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    out = _base_line(args, "positions/s per ePIE iteration", n / (ms * 1e-3), "positions/s", ms, "c64",
                     "synthetic (seeded random-field object, Gaussian probe, Poisson 1e6 photons/frame)",
                     {"workload": "ePIE 12x12 raster, 128x128 frames, 320x320 object, 1 probe mode, probe update on",
                      "frames": n, "frame": m})
    out["us_per_position"] = ms * 1e3 / n
    out["roofline"] = None
    if not args.no_cpu:
        import oracle
        sc = e.get_scalars()
        st = oracle.EpieState(e.get_object(), e.get_probe(), sc[E.SCAL_OBJ_MAX], sc[E.SCAL_PROBE_MAX])
        a64 = amp_deq.astype(np.float64)
        iters, t1 = 0, time.perf_counter()
        while True:
            oracle.epie_iteration(st, a64, W["pos"], rng.permutation(n), update_probe=True)
            iters += 1
            dt = time.perf_counter() - t1
            if dt >= 10.0 and iters >= 2:
                break
        out["cpu_baseline"] = {"value": n * iters / dt, "unit": "positions/s", "cores": oracle.num_threads(),
                               "kind": "port", "sample": f"{iters} FP64 oracle ePIE iterations over all {n} positions; "
                                                         f"{dt:.1f} s on {cpu_model()}", **host_info()}
    e.close()
    return out


def run_batch(args) -> dict:
    """16 independent DM scans of uneven size advanced together: one captured HIP graph per flag
    set, replayed per iteration; the same scans run one after another beside it."""
    import torch
    from ptychopy_amd import workloads, autotune
    from ptychopy_amd import _engine as E
    scans = workloads.batch_c3()
    tot = sum(s["n"] for s in scans)
    waves = [autotune.tune_dm_wave(s["pos"], s["m"], s["canvas"], iters=2)[0] for s in scans]

    def make():
        es = []
        for s, w in zip(scans, waves):
            e = E.Engine(s["n"], s["m"], s["canvas"])
            e.set_frames(s["amp"])
            e.set_positions(s["pos"])
            e.clear_object()
            e.init_probe(s["m"] * 24.0 / 128.0 * 1e-8, 1e-8)
            e.set_tuning(E.TUNE_DM_WAVE, w)
            es.append(e)
        return es

    def flags(i):
        return (E.FLAG_PHASE_CONSTRAINT if i <= 1 else 0) | (E.FLAG_UPDATE_PROBE if i >= 10 else 0)

    es = make()
    with E.Batch(es) as b:
        it = 0
        for _ in range(max(args.warmup, 10)):
            it += 1
            b.dm_iteration(flags(it))
        b.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            it += 1
            b.dm_iteration(flags(it))
        b.synchronize()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        graphs = b.graph_count
    for e in es:
        e.close()
    es = make()
    it = 0
    for _ in range(max(args.warmup, 10)):
        it += 1
        for e in es:
            e.dm_iteration(flags(it))
    for e in es:
        e.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        it += 1
        for e in es:
            e.dm_iteration(flags(it))
    for e in es:
        e.synchronize()
    ms_seq = (time.perf_counter() - t0) * 1e3 / args.steps
    for e in es:
        e.close()
    out = _base_line(args, "diffraction frames/s per iteration", tot / (ms * 1e-3), "frames/s", ms, "c64",
                     "synthetic (16 seeded scans: random-field objects, Gaussian probes, Poisson 1e5-1e7 photons)",
                     {"workload": "batched DM, 16 scans of 20x20..51x51 positions, frames 128/256/512, "
                                  "one captured HIP graph per iteration", "frames": tot,
                      "scans": [[s["n"], s["m"]] for s in scans], "frames_per_wave": waves})
    out["sequential_ms_per_step"] = ms_seq
    out["graphs"] = graphs
    out["roofline"] = None
    return out


def run_mls(args) -> dict:
    import torch
    from ptychopy_amd import workloads, _driver
    from ptychopy_amd import _engine as E

    t_gen = time.perf_counter()
    W = workloads.mls_c2()
    t_gen = time.perf_counter() - t_gen
    n, m, k = W["n"], W["m"], W["k"]
    rows, cols = W["canvas"]
    rng = np.random.default_rng(20261204)
    from ptychopy_amd import autotune
    if args.mls_shape:
        cb, fpc, graphs = (int(v) for v in args.mls_shape.split(","))
        tune_ms = None
    else:
        (cb, fpc, graphs), tune_ms = autotune.tune_mls(W["pos"], m, k, (rows, cols), W["box"])
    e = E.Engine(n, m, (rows, cols), n_modes=k)
    e.set_tuning(E.TUNE_MLS_MOD_CB, cb)
    e.set_tuning(E.TUNE_MLS_PUM_FPC, fpc)
    e.set_tuning(E.TUNE_MLS_GRAPHS, graphs)
    e.set_frames(W["amp"])
    amp_deq = e.get_frames()
    obj0 = np.ones((rows, cols), np.complex64)
    e.set_object(obj0)
    e.set_probe(W["probe_true"])
    vmode = (rng.uniform(0, 1, (m, m)) * np.exp(2j * np.pi * rng.uniform(0, 1, (m, m)))).astype(np.complex64)
    evo1 = 1e-6 * rng.standard_normal(n)
    fp = (amp_deq.astype(np.float64) ** 2).sum(axis=(1, 2))
    cfg = E.mls_config(W["box"])
    e.mls_begin(cfg, W["pos"], vmode, evo1, fp)
    pps = int(cfg.probe_pos_search)
    parts = _driver.mls_partitions(n, 100, 1)
    n_groups = len(parts[0])

    def groups(i):
        p = parts[i % len(parts)]
        return [p[j] for j in rng.permutation(len(p))]

    it = 0
    for _ in range(args.warmup):
        it += 1
        e.mls_iteration(it, groups(it))
    # timed iterations are numbered from the position-search iteration on: every one runs the
    # position refinement, the variable-probe and intensity updates and no Fourier errors
    it = max(it, pps - 1)
    e.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        it += 1
        e.mls_iteration(it, groups(it))
    e.synchronize()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # per-kind launch times from HIP events on the engine stream (extra iterations of the same kind)
    kt = {}
    exit_launches = 0.0
    for _ in range(args.kernel_iters):
        it += 1
        gl = groups(it)
        for key, v in e.mls_iteration_timed(it, gl).items():
            kt[key] = kt.get(key, 0.0) + v / args.kernel_iters
        # the exit-wave row pass runs as two launches (modes 2..K ahead of the side streams' join,
        # then mode 0) for every group after the first that is no smaller than the one before it
        sizes = [len(x) for x in gl]
        exit_launches += (len(gl) + sum(1 for a, b in zip(sizes, sizes[1:]) if b >= a)) / args.kernel_iters
    # the dominant one-kernel kind (the two-kernel probe-mode update is timed as a pair), by the
    # kernels' own time: start/stop events recorded by each launch's dispatch (the marks' intervals
    # in kernel_ms also carry the events' own processing on the stream)
    launches = {key: (exit_launches if key == "exit_row_pass" else n_groups)
                for key in ("exit_row_pass", "modulus_column_pass", "inverse_row_pass")}
    own = {key: max(kt[key + "_own"], 1e-6) for key in launches}
    dom = max(own, key=own.get)
    dom_bytes = workloads.mls_kind_bytes(dom, m, k, n, n_groups)
    ach = dom_bytes / (own[dom] * 1e-3) / 1e9
    traffic = None
    tfile = ROOT / "profiles" / "mls_c2_traffic.json"
    if tfile.exists():
        traffic = json.loads(tfile.read_text()).get("kinds", {}).get(dom, {}).get("traffic")
    bpf = workloads.mls_bytes_per_frame(m, rows * cols, n, n_groups)
    fpf = workloads.mls_flops_per_frame(m, k)
    gbs = bpf * n / (ms * 1e-3) / 1e9
    tfs = fpf * n / (ms * 1e-3) / 1e12
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
        "data": "synthetic (seeded random-field object, 5 Hermite-Gaussian modes, Poisson 1e5 photons/frame)",
        "config": {"workload": "MLs 51x51 raster, 256x256 frames, 5 probe modes, FP16 amplitudes, "
                               f"{n_groups} groups, probe/object/variable-probe updates and position refinement "
                               "in every timed iteration",
                   "frames": n, "frame": m, "canvas": [rows, cols], "modes": k, "groups": n_groups,
                   "timed_iterations": [it - args.kernel_iters - args.steps + 1, it - args.kernel_iters],
                   "position_refinement": True, "fourier_errors": False},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": traffic, "algorithmic_bytes": dom_bytes,
                     "launches": launches[dom], "ms_per_iteration": own[dom]},
        "iteration_roofline": {"bytes_per_frame": bpf, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": gbs / HBM_PEAK_GBS},
        "flop_roofline": {"flops_per_frame": fpf, "achieved": tfs, "peak": 157.3, "unit": "TFLOP/s",
                          "frac": tfs / 157.3},
        "kernel_ms": kt,
        "tuning": {"mls_mod_cb": cb, "mls_pum_fpc": fpc, "mls_group_graphs": graphs, "table_ms": tune_ms},
        "setup_s": t_gen,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = mls_cpu_baseline(W, amp_deq, W["probe_true"], obj0, vmode, evo1, groups(it + 1),
                                               max(it + 1, pps + 1))
    e.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("mls_c2", "dm_c1", "epie_c0", "batch_c3", "c4"), default="mls_c2")
    ap.add_argument("--c4-train", type=int, default=50000, help="c4: training frames")
    ap.add_argument("--c4-held", type=int, default=5000, help="c4: held-out frames")
    ap.add_argument("--c4-epochs", type=int, default=2, help="c4: training epochs")
    ap.add_argument("--c4-target", type=float, default=1e-2, help="c4: normalised object error target")
    ap.add_argument("--c4-iters", type=int, default=100, help="c4: MLs iteration limit per start")
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 20 for dm_c1, 10 for mls_c2)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default 10 for dm_c1, 5 for mls_c2)")
    ap.add_argument("--cpu-frames", type=int, default=900)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-iters", type=int, default=5)
    ap.add_argument("--copy-probe", action="store_true",
                    help="first copy 512 MiB device-to-device twice from a cache-cold source, with the runtime's "
                         "copy kernel and with an elementwise kernel (kernels of known traffic that scale the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiled run)")
    ap.add_argument("--mls-shape", default=None,
                    help="MLs launch shape 'CB,FPC,GRAPHS' (modulus block width, probe-mode chunk frames, "
                         "group graphs 0/1); default: the autotuner's choice")
    ap.add_argument("--dm-graph", type=int, default=0, help="dm_c1: replay iterations from captured HIP graphs")
    ap.add_argument("--dm-wave", type=int, default=None,
                    help="frames per DM wave (default: the autotuner's choice; 0 automatic, -1 one pass)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} / WORLD_SIZE {world} is not supported "
              "(the DM frame loop would shard by scan region with one object all-reduce per iteration; "
              "that path is not built)", file=sys.stderr)
        return 2
    if args.copy_probe:
        import torch
        src = torch.ones(128 << 20, dtype=torch.float32, device="cuda")
        evict = torch.ones(256 << 20, dtype=torch.float32, device="cuda")  # 1 GiB: src leaves the caches
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)  # the runtime's copy kernel
        evict.fill_(2.0)
        torch.mul(src, 1.0, out=dst)  # an elementwise kernel with 16-byte loads and stores
        torch.cuda.synchronize()
        del src, dst, evict
    if args.workload == "c4":
        from ptychopy_amd import warmstart
        t0 = time.perf_counter()
        r = warmstart.c4_experiment(n_train=args.c4_train, n_held=args.c4_held, epochs=args.c4_epochs,
                                    target=args.c4_target, max_iter=args.c4_iters,
                                    log=lambda s: print(s, file=sys.stderr, flush=True))
        out = {"metric": f"time-to-{args.c4_target:g} normalised object error, warm vs plain start",
               "value": r["speedup"], "unit": "x (plain time / warm time)", "n_gpus": 1, "steps": None,
               "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "bf16 network, c64 solver",
               "data": "synthetic (random-field training objects, phase std 0.2-1.0 rad; the MLs workload's "
                       "seeded object and Poisson frames)",
               "config": {"workload": "learned warm start + MLs, 51x51 raster, 256x256 frames, 5 probe modes"},
               "roofline": {"bound": "mfma", "kernel": "network training step (all convolutions)",
                            "achieved": r["train_tflops"], "peak": 2500.0, "unit": "TFLOP/s",
                            "frac": r["train_tflops"] / 2500.0, "traffic": None},
               "c4": r, "wall_s": time.perf_counter() - t0}
        print(json.dumps(out))
        return 0
    if args.workload in ("epie_c0", "batch_c3"):
        args.steps = 20 if args.steps is None else args.steps
        args.warmup = 10 if args.warmup is None else args.warmup
        print(json.dumps(run_epie(args) if args.workload == "epie_c0" else run_batch(args)))
        return 0
    if args.workload == "mls_c2":
        args.steps = 10 if args.steps is None else args.steps
        args.warmup = 5 if args.warmup is None else args.warmup
        print(json.dumps(run_mls(args)))
        return 0
    args.steps = 20 if args.steps is None else args.steps
    args.warmup = 10 if args.warmup is None else args.warmup

    import torch
    from ptychopy_amd import workloads
    from ptychopy_amd._engine import Engine

    t_gen = time.perf_counter()
    wl = workloads.dm_c1()
    t_gen = time.perf_counter() - t_gen
    n, m = wl["n"], wl["m"]
    rows, cols = wl["canvas"]
    from ptychopy_amd._engine import TUNE_DM_WAVE
    tune_ms = None
    if args.dm_wave is None:
        from ptychopy_amd import autotune
        args.dm_wave, tune_ms = autotune.tune_dm_wave(wl["pos"], m, (rows, cols))
    e = Engine(n, m, (rows, cols))
    e.set_tuning(TUNE_DM_WAVE, args.dm_wave)
    if args.dm_graph:
        from ptychopy_amd._engine import TUNE_DM_GRAPH
        e.set_tuning(TUNE_DM_GRAPH, 1)
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
        e.dm_iteration(dm_flags(it, timed=True))
    e.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps

    # per-kernel durations from HIP events on the engine stream
    kt = np.zeros(len(KERNELS))
    for _ in range(args.kernel_iters):
        it += 1
        kt += e.dm_iteration_timed(dm_flags(it, timed=True))
    kt /= args.kernel_iters
    obj_px = rows * cols
    dom = int(np.argmax(kt))
    name = KERNELS[dom]
    ach = kernel_bytes(name, n, m, obj_px) / (kt[dom] * 1e-3) / 1e9
    # HBM bytes of this kernel kind per iteration, from the committed FETCH_SIZE / WRITE_SIZE counter
    # runs (scaled by the calibration copy), valid for the default 64-frame waves
    traffic = None
    tfile = ROOT / "profiles" / "dm_c1_traffic.json"
    if tfile.exists() and args.dm_wave == 64:
        traffic = json.loads(tfile.read_text()).get("kinds", {}).get(name, {}).get("traffic")
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
                               "1 probe mode, FP16 amplitudes, probe update on in every timed iteration",
                   "frames": n, "frame": m, "canvas": [rows, cols], "modes": 1, "probe_update": True,
                   "phase_constraint": False},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": kernel_bytes(name, n, m, obj_px)},
        "iteration_roofline": {"bytes_per_frame": bpf, "achieved": it_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": it_gbs / HBM_PEAK_GBS},
        "kernel_ms": {k: float(v) for k, v in zip(KERNELS, kt)},
        "dm_wave_frames": e.get_tuning(1),
        "tuning": {"dm_wave": args.dm_wave, "dm_graph": args.dm_graph, "table_ms": tune_ms},
        "setup_s": t_gen,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_frames)
    e.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
