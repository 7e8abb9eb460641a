"""Autotuner host logic (no GPU): a persisted timing table is read back instead of timing again,
keyed by the problem shape, and the fastest entry is chosen."""
import json

import numpy as np

from ptychopy_amd import autotune


def test_dm_wave_table_hit_returns_the_fastest(tmp_path):
    path = tmp_path / "t.json"
    pos = np.zeros((900, 2), np.float32)
    key = autotune.dm_key(900, 512, (1027, 1027))
    path.write_text(json.dumps({key: {"-1": 5.5, "32": 3.2, "64": 2.9, "128": 2.95}}))
    best, times = autotune.tune_dm_wave(pos, 512, (1027, 1027), path=path)
    assert best == 64 and times[64] == 2.9 and set(times) == {-1, 32, 64, 128}


def test_mls_table_hit_returns_the_fastest_shape(tmp_path):
    path = tmp_path / "t.json"
    pos = np.zeros((2601, 2), np.float32)
    key = autotune.mls_key(2601, 256, 5, (1146, 1146), 26)
    table = {"8,2,0": 33.0, "16,8,0": 30.2, "32,4,0": 31.5, "16,8,1": 29.9}
    path.write_text(json.dumps({key: table, "other": {"1,1,1": 0.1}}))
    best, times = autotune.tune_mls(pos, 256, 5, (1146, 1146), (1, 2, 1, 2), path=path)
    assert best == (16, 8, 1) and times == table
    # another shape misses the table (a different key)
    assert autotune.mls_key(2600, 256, 5, (1146, 1146), 26) != key
