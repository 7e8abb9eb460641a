"""Batched multi-scan DM (captured HIP graphs): each scan equals its standalone run bitwise."""
import numpy as np
import pytest

from problems import make_problem

pytestmark = pytest.mark.gpu

SCANS = [dict(m=32, n_side=4, seed=11), dict(m=64, n_side=5, seed=12), dict(m=128, n_side=3, seed=13),
         dict(m=128, n_side=4, seed=14)]


def _engine(P):
    from ptychopy_amd._engine import Engine
    e = Engine(P["n"], P["m"], P["shape"])
    e.set_frames(P["amp"])
    e.set_positions(P["pos"])
    e.clear_object()
    e.init_probe(1e-7, 1e-8)
    return e


def _flags(i):
    from ptychopy_amd._engine import FLAG_PHASE_CONSTRAINT, FLAG_UPDATE_PROBE
    return (FLAG_PHASE_CONSTRAINT if i <= 1 else 0) | (FLAG_UPDATE_PROBE if i >= 10 else 0)


def test_batch_equals_standalone_runs_bitwise(gpu):
    from ptychopy_amd._engine import Batch
    probs = [make_problem(**kw) for kw in SCANS]
    alone = []
    for P in probs:
        e = _engine(P)
        for i in range(1, 13):
            e.dm_iteration(_flags(i))
        alone.append((e.get_object(), e.get_probe(), e.get_psi()))
        e.close()
    engines = [_engine(P) for P in probs]
    with Batch(engines) as b:
        for i in range(1, 13):
            b.dm_iteration(_flags(i))
        b.synchronize()
        # iteration 1 runs eagerly (first-call factors); then one graph per flag set
        assert b.graph_count == 2
        for e, (o, p, s) in zip(engines, alone):
            assert np.array_equal(e.get_object(), o)
            assert np.array_equal(e.get_probe(), p)
            assert np.array_equal(e.get_psi(), s)
    for e in engines:
        e.close()


def test_batch_recaptures_after_a_scan_changes(gpu):
    from ptychopy_amd._engine import Batch, TUNE_DM_WAVE
    P = make_problem(m=128, n_side=4, seed=15)

    def standalone(switch):
        e = _engine(P)
        for i in range(1, 6):
            if switch and i == 4:
                e.set_tuning(TUNE_DM_WAVE, 4)
            e.dm_iteration(0)
        out = e.get_object()
        e.close()
        return out

    ref1, ref2 = standalone(False), standalone(True)
    e1, e2 = _engine(P), _engine(P)
    with Batch([e1, e2]) as b:
        for i in range(1, 4):
            b.dm_iteration(0)
        e2.set_tuning(TUNE_DM_WAVE, 4)  # new launch shape for scan 2: the captured graph is stale
        for i in range(4, 6):
            b.dm_iteration(0)
        b.synchronize()
        assert np.array_equal(e1.get_object(), ref1)
        assert np.array_equal(e2.get_object(), ref2)
    e1.close()
    e2.close()


def test_autotuner_persists_a_timing_table(gpu, tmp_path):
    import json
    from ptychopy_amd import autotune
    P = make_problem(m=128, n_side=4, seed=16)
    path = tmp_path / "tune.json"
    best, times = autotune.tune_dm_wave(P["pos"], 128, P["shape"], iters=1, path=path)
    assert best in times and set(times) <= set(autotune.DM_WAVE_CANDIDATES) and -1 in times
    table = json.loads(path.read_text())
    key = autotune.dm_key(P["n"], 128, P["shape"])
    assert set(int(c) for c in table[key]) == set(times)
    # a second call reads the table instead of timing again
    assert autotune.tune_dm_wave(P["pos"], 128, P["shape"], path=path) == (best, times)


def test_batch_replays_after_positions_move(gpu):
    """Moving one scan's positions between batch iterations rebuilds its frame lists (often at the
    same device addresses); the replayed iterations must use the new row ranges: bitwise equal to a
    standalone run with the same move."""
    from ptychopy_amd._engine import Batch
    P = make_problem(m=128, n_side=4, seed=17)
    moved = P["pos"] + np.float32(1.0)
    assert (np.ceil(moved.max(axis=0)) + 128 <= np.array(P["shape"])).all()

    def standalone(move):
        e = _engine(P)
        for i in range(1, 7):
            if move and i == 4:
                e.set_positions(moved)
            e.dm_iteration(0)
        out = e.get_object(), e.get_psi()
        e.close()
        return out

    ref_still, ref_moved = standalone(False), standalone(True)
    e1, e2 = _engine(P), _engine(P)
    with Batch([e1, e2]) as b:
        for i in range(1, 4):
            b.dm_iteration(0)
        e2.set_positions(moved)
        for i in range(4, 7):
            b.dm_iteration(0)
        b.synchronize()
        assert np.array_equal(e1.get_object(), ref_still[0]) and np.array_equal(e1.get_psi(), ref_still[1])
        assert np.array_equal(e2.get_object(), ref_moved[0]) and np.array_equal(e2.get_psi(), ref_moved[1])
    e1.close()
    e2.close()


@pytest.mark.parametrize("m,wave", [(128, 0), (128, -1), (32, 0)])
def test_standalone_dm_graph_replays_equal_direct_launches(gpu, m, wave):
    """Standalone DM iterations replayed from graphs captured on the handle's stream (one per flag
    set: phase constraint, plain, probe update; re-captured after the positions move, which
    rebuilds the frame lists) equal direct launches bit for bit, with host reads of the state
    between the replays."""
    from ptychopy_amd._engine import TUNE_DM_GRAPH, TUNE_DM_WAVE, FLAG_PHASE_CONSTRAINT, FLAG_UPDATE_PROBE
    P = make_problem(m=m, n_side=4, seed=41)
    moved = P["pos"] + np.float32(1.0)
    assert (np.ceil(moved.max(axis=0)) + m <= np.array(P["shape"])).all()
    runs = []
    for graph in (1, 0):
        e = _engine(P)
        e.set_tuning(TUNE_DM_WAVE, wave)
        e.set_tuning(TUNE_DM_GRAPH, graph)
        assert e.get_tuning(TUNE_DM_GRAPH) == graph
        snaps = []
        for i in range(1, 10):
            if i == 6:
                e.set_positions(moved)
            e.dm_iteration((FLAG_PHASE_CONSTRAINT if i <= 1 else 0) | (FLAG_UPDATE_PROBE if i >= 4 else 0))
            snaps.append(e.get_object())
        runs.append((snaps, e.get_probe(), e.get_psi()))
        e.close()
    for i, (a, b) in enumerate(zip(runs[0][0], runs[1][0])):
        assert np.array_equal(a, b), f"object after iteration {i + 1}"
    assert np.array_equal(runs[0][1], runs[1][1]) and np.array_equal(runs[0][2], runs[1][2])
