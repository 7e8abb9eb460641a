"""Autotuner for the engine's launch-shape knobs (no reference counterpart).

``tune_dm_wave`` times one DM iteration for each candidate frames-per-wave value of
``PTYK_TUNE_DM_WAVE``, and ``tune_mls`` one MLs iteration for each pair of the modulus pass's
block width (``PTYK_TUNE_MLS_MOD_CB``) and the probe-mode update's frames per chunk
(``PTYK_TUNE_MLS_PUM_FPC``), on a scratch handle with the scan's own geometry (the timing does
not depend on the data values).  Each keeps the fastest and persists its timing table as JSON
(``PTYK_TUNE_TABLE`` or ``ptychopy_amd/_tune/tuning.json``) keyed by the problem shape.  The
choice changes only the schedule: every candidate computes the same iteration (the wave and
single-pass DM forms agree to FP32 rounding, the MLs block widths bit for bit, the chunk sizes to
FP32 rounding of the chunk sums; see the parity tests).
"""
from __future__ import annotations

import json
import os
import time
from pathlib import Path

import numpy as np

from ._engine import Engine, TUNE_DM_WAVE

DM_WAVE_CANDIDATES = (-1, 16, 32, 64, 128, 256)


def table_path() -> Path:
    p = os.environ.get("PTYK_TUNE_TABLE")
    return Path(p) if p else Path(__file__).resolve().parent / "_tune" / "tuning.json"


def _load(path: Path) -> dict:
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return {}


def _store(path: Path, table: dict):
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_suffix(".tmp")
    tmp.write_text(json.dumps(table, indent=1, sort_keys=True))
    tmp.replace(path)


def dm_key(n: int, m: int, canvas, k: int = 1) -> str:
    return f"dm_wave/M{m}/N{n}/K{k}/{int(canvas[0])}x{int(canvas[1])}"


def tune_dm_wave(pos, m: int, canvas, k: int = 1, iters: int = 3, candidates=DM_WAVE_CANDIDATES,
                 path: Path | None = None, force: bool = False, device: int = 0):
    """Best frames-per-wave for a DM problem with these positions; returns (best, {candidate: ms})."""
    pos = np.ascontiguousarray(pos, np.float32)
    n = pos.shape[0]
    path = path or table_path()
    table = _load(path)
    key = dm_key(n, m, canvas, k)
    if not force and key in table:
        t = {int(c): v for c, v in table[key].items()}
        return min(t, key=t.get), t
    cands = [c for c in candidates if c == -1 or (m >= 128 and c <= n)]
    times = {}
    with Engine(n, m, canvas, n_modes=k, device=device) as e:
        e.set_frames(np.ones((n, m, m), np.float32))
        e.set_positions(pos)
        e.set_object(np.ones(canvas, np.complex64))
        e.set_probe(np.ones((k, m, m), np.complex64) / m)
        e.refresh_maxima()
        for c in cands:
            e.set_tuning(TUNE_DM_WAVE, c)
            e.dm_iteration(0)  # builds the schedule's frame lists
            e.synchronize()
            ts = []
            for _ in range(iters):
                t0 = time.perf_counter()
                e.dm_iteration(0)
                e.synchronize()
                ts.append(time.perf_counter() - t0)
            times[c] = float(np.median(ts) * 1e3)
    table[key] = {str(c): v for c, v in times.items()}
    _store(path, table)
    return min(times, key=times.get), times


MLS_CB_CANDIDATES = (8, 16, 32)
MLS_FPC_CANDIDATES = (2, 4, 8)


def mls_key(n: int, m: int, k: int, canvas, group: int) -> str:
    return f"mls/M{m}/N{n}/K{k}/G{group}/{int(canvas[0])}x{int(canvas[1])}"


def tune_mls(pos, m: int, k: int, canvas, box, iters: int = 2, cb_candidates=MLS_CB_CANDIDATES,
             fpc_candidates=MLS_FPC_CANDIDATES, path: Path | None = None, force: bool = False, device: int = 0):
    """Best (modulus block width, probe-mode chunk frames, group graphs on/off) for an MLs problem
    with these positions, timed on a scratch handle over full iterations (position refinement and
    the variable probe on, the reference's max(1, N/100)-frame groups): every (cb, fpc) pair with
    direct launches, then the best pair replayed from group graphs.  Returns ((cb, fpc, graphs),
    {"cb,fpc,graphs": ms})."""
    from . import _driver
    from ._engine import TUNE_MLS_GRAPHS, TUNE_MLS_MOD_CB, TUNE_MLS_PUM_FPC, mls_config
    pos = np.ascontiguousarray(pos, np.float32)
    n = pos.shape[0]
    path = path or table_path()
    table = _load(path)
    group = max(1, n // 100)
    key = mls_key(n, m, k, canvas, group)
    if not force and key in table:
        t = table[key]
        best = min(t, key=t.get)
        return tuple(int(v) for v in best.split(",")), t
    rng = np.random.default_rng(0)
    times = {}
    with Engine(n, m, canvas, n_modes=k, device=device) as e:
        e.set_frames(rng.uniform(0.5, 1.5, (n, m, m)).astype(np.float32))
        e.set_object(np.ones(canvas, np.complex64))
        yy, xx = np.mgrid[0:m, 0:m] - m / 2
        g = np.exp(-(xx ** 2 + yy ** 2) / (2 * (m / 10) ** 2))
        e.set_probe(np.stack([g * (0.5 ** i) * np.exp(1j * i * xx / m) for i in range(k)]).astype(np.complex64))
        vmode = (rng.uniform(0, 1, (m, m)) * np.exp(2j * np.pi * rng.uniform(0, 1, (m, m)))).astype(np.complex64)
        cfg = mls_config(box, probe_pos_search=1)
        e.mls_begin(cfg, pos, vmode, 1e-6 * rng.standard_normal(n))
        part = _driver.mls_partitions(n, 1, 0)[0]
        it = 3  # position refinement, variable probe and intensity all on

        def timed(cb, fpc, graphs):
            e.set_tuning(TUNE_MLS_MOD_CB, cb)
            e.set_tuning(TUNE_MLS_PUM_FPC, fpc)
            e.set_tuning(TUNE_MLS_GRAPHS, graphs)
            e.mls_iteration(it, part)  # warm-up of this shape (and the graphs' capture)
            ts = []
            for _ in range(iters):
                t0 = time.perf_counter()
                e.mls_iteration(it, part)
                ts.append(time.perf_counter() - t0)
            times[f"{cb},{fpc},{graphs}"] = float(np.median(ts) * 1e3)

        for cb in cb_candidates:
            for fpc in fpc_candidates:
                timed(cb, fpc, 0)
        cb, fpc, _ = (int(v) for v in min(times, key=times.get).split(","))
        timed(cb, fpc, 1)
    table[key] = times
    _store(path, table)
    return tuple(int(v) for v in min(times, key=times.get).split(",")), times
