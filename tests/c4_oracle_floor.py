"""The FP64 oracle's MLs error curve e(t) on the 51 x 51 / 256^2 / 5-mode workload (or a smaller
raster of the same shape), CPU only: the configuration's floor, independent of the GPU path.

Data are formed on the CPU with the oracle's simulator (the workload's seeded object, the five
Hermite-Gaussian modes, 14 px raster steps, 1e5 photons per frame with Poisson noise); the solver
starts from a uniform object and the true probe modes, as the learned-warm-start experiment's
plain start does, with the reference's max(1, N/100)-frame groups (26 frames at 51 x 51) drawn
afresh each iteration.  Prints one line per iteration:

    python tests/c4_oracle_floor.py [--side 51] [--iters 100]
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def main() -> int:
    import oracle
    from ptychopy_amd import sim
    from ptychopy_amd._driver import cartesian_mesh_raw, mls_canvas
    ap = argparse.ArgumentParser()
    ap.add_argument("--side", type=int, default=51)
    ap.add_argument("--iters", type=int, default=100)
    ap.add_argument("--group", type=int, default=0, help="frames per group (0: max(1, N/100))")
    args = ap.parse_args()
    m, k, photons = 256, 5, 1e5
    dx_s = 100e-9 / 14.0
    raw = cartesian_mesh_raw(args.side, args.side, 100e-9, 100e-9, dx_s, m)
    lo, hi = np.minimum(raw.min(axis=0), 0), np.maximum(raw.max(axis=0), 0)
    canvas, box = mls_canvas(raw, np.array([lo[0], lo[1], hi[0], hi[1]], np.float32), m)
    obj = sim.random_field_object(canvas, sim.rng_for(2, 0), phase_std=0.5, phase_corr=6.0)
    modes = sim.hermite_gaussian_modes(m, k, 48.0)
    off = np.array([np.ceil(canvas[0] / 2 - m / 2) - 1, np.ceil(canvas[1] / 2 - m / 2) - 1], np.float32)
    t0 = time.perf_counter()
    inten = sum(oracle.simulate(obj, modes[i], raw + off, squared=True) for i in range(k))
    amp = sim.poisson_amplitudes(inten, photons, np.random.default_rng(7)).astype(np.float64)
    del inten
    n = len(raw)
    rng = np.random.default_rng(0)
    vmode = (rng.uniform(0, 1, (m, m)) * np.exp(2j * np.pi * rng.uniform(0, 1, (m, m)))).astype(np.complex64)
    evo1 = 1e-6 * rng.standard_normal(n)
    st = oracle.MlsState(np.ones(canvas, np.complex128), modes, vmode, evo1, raw, amp)
    st.init((amp ** 2).sum(axis=(1, 2)))
    g = args.group or max(1, n // 100)
    print(f"# {n} frames of {m}^2, {k} modes, canvas {canvas}, groups of {g}, {oracle.num_threads()} threads; "
          f"data in {time.perf_counter() - t0:.1f} s", flush=True)
    print(f"iter 0 e {oracle.object_error(st.obj, obj, st.probe, st.roi)[0]:.5f}", flush=True)
    grng = np.random.default_rng(1)
    t0 = time.perf_counter()
    for it in range(1, args.iters + 1):
        idx = grng.permutation(n).astype(np.int32)
        st.iteration(oracle.mls_params(box, iter=it), [idx[i:i + g] for i in range(0, n, g)])
        e = oracle.object_error(st.obj, obj, st.probe, st.roi)[0]
        print(f"iter {it} e {e:.5f} t {time.perf_counter() - t0:.0f} s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
