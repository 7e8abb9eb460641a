"""Autotuner for the engine's launch-shape knobs (no reference counterpart).

``tune_dm_wave`` times one DM iteration for each candidate frames-per-wave value of
``PTYK_TUNE_DM_WAVE`` on a scratch handle with the scan's own geometry (the timing does not
depend on the data values), keeps the fastest, and persists the timing table as JSON
(``PTYK_TUNE_TABLE`` or ``ptychopy_amd/_tune/tuning.json``) keyed by the problem shape.  The
choice changes only the schedule: every candidate computes the same iteration (the wave and
single-pass forms agree to FP32 rounding; see the DM parity tests).
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
