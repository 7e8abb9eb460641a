"""MLs on the GPU through the C ABI against the FP64 oracle (oracle/oracle_mls.cpp).

Each iteration check starts the oracle from the GPU's own state (object, probe modes, variable
mode, evolution, illumination, positions, ROIs, steps), runs one iteration on both with the same
groups, and compares: object and probe within 1e-5 relative (L2), ROI origins bit-exact."""
import numpy as np
import pytest

from conftest import rel_err
from problems import make_mls_problem, mls_groups

pytestmark = pytest.mark.gpu


def _engine_for(P, amp, **cfg_kw):
    from ptychopy_amd import _engine
    st = P["state"]
    k, m, _ = P["probe0"].shape
    e = _engine.Engine(P["n"], m, P["shape"], n_modes=k)
    e.set_frames(amp)
    e.set_object(P["obj0"])
    e.set_probe(P["probe0"])
    cfg = _engine.mls_config(P["box"], **cfg_kw)
    e.mls_begin(cfg, P["pos"], P["vmode0"], P["evo1_0"], P["frame_power"])
    return e


def _oracle_from_engine(e, orc, amp):
    g = e.mls_get_state()
    o = orc.MlsState(e.get_object(), e.get_probe(), g["variable_mode"], g["evolution"][1], g["positions_start"],
                     amp)
    o.pos = g["positions"].copy()
    o.evo = g["evolution"].astype(np.float64).copy()
    o.illum = g["illumination"].astype(np.float64)
    o.max_illum = float(g["max_illumination"])
    o.roi = g["roi"].copy()
    o.sub = g["subpixel"].copy()
    o.beta_obj = g["beta_object"].astype(np.complex128)
    o.beta_probe = g["beta_probe"].astype(np.complex128)
    return o


def _problem(k=2, m=32, n_side=5, seed=0):
    # evolution draws all non-zero: a group whose draws are all zero divides by zero in the
    # variable-probe update (Probe.cpp:1387-1441), in the reference as here
    P = make_mls_problem(m=m, n_side=n_side, k=k, seed=seed, phase_object=False, photons=2e5, zero_evo01=False)
    return P


def test_mls_begin_matches_oracle_init(gpu, orc):
    P = _problem(k=3)
    e = _engine_for(P, P["state"].amp)
    amp = e.get_frames().astype(np.float64)
    o = orc.MlsState(P["obj0"], P["probe0"], P["vmode0"], P["evo1_0"], P["pos"], amp)
    o.init(P["frame_power"])
    g = e.mls_get_state()
    assert np.array_equal(g["roi"], o.roi)
    assert np.array_equal(g["subpixel"], o.sub)
    assert rel_err(e.get_object(), o.obj) < 1e-6
    assert rel_err(e.get_probe(), o.probe) < 1e-6
    assert rel_err(g["variable_mode"], o.vmode) < 1e-6
    assert np.allclose(g["evolution"], o.evo, rtol=1e-6, atol=0)
    assert rel_err(g["illumination"], o.illum) < 1e-6
    assert abs(g["max_illumination"] / o.max_illum - 1) < 1e-6
    assert np.all(g["beta_object"] == 1) and np.all(g["beta_probe"] == 1)


CONFIGS = [
    dict(k=2, cfg=dict(probe_pos_search=3)),
    dict(k=1, cfg=dict(probe_pos_search=2, variable_probe=0)),
    dict(k=3, cfg=dict(apply_subpix_shift=0, probe_pos_search=99)),
    dict(k=2, cfg=dict(variable_probe=0, variable_intensity=0, variable_probe_modes=0, beta_lsq=0.0)),
    dict(k=2, cfg=dict(apply_multimodal_update=1, probe_pos_search=4)),
    dict(k=5, cfg=dict(probe_pos_search=3)),  # five modes: the fused update of modes 2..5 (8-wide form)
    # 128^2 / 256^2 frames take the three-pass modulus (row pass, K-mode 64 x 32 blocks, row pass)
    dict(k=1, m=128, n_side=3, cfg=dict(probe_pos_search=2, variable_probe=0)),
    dict(k=3, m=128, n_side=3, cfg=dict(probe_pos_search=3)),
    dict(k=5, m=256, n_side=3, cfg=dict(probe_pos_search=99)),
    # position refinement at 256^2 and 512^2 (derivative spectra formed in the inverse row pass)
    dict(k=2, m=256, n_side=3, cfg=dict(probe_pos_search=2)),
    dict(k=1, m=512, n_side=3, cfg=dict(probe_pos_search=2, variable_probe=0)),
    # groups of 9 and 7 frames: the probe-mode update of modes 2..K sums over several chunks
    dict(k=3, m=128, n_side=4, group=9, cfg=dict(probe_pos_search=2)),
    # multimodal object updates at 128^2: every mode's own inverse pass and object step, no
    # second-stream windows for the variable mode
    dict(k=3, m=128, n_side=3, cfg=dict(apply_multimodal_update=1, probe_pos_search=2)),
]


@pytest.mark.parametrize("case", range(len(CONFIGS)))
def test_mls_iterations_match_oracle(gpu, orc, case):
    c = CONFIGS[case]
    P = _problem(k=c["k"], m=c.get("m", 32), n_side=c.get("n_side", 5), seed=case)
    e = _engine_for(P, P["state"].amp, **c["cfg"])
    amp = e.get_frames().astype(np.float64)
    okw = {("pps" if key == "probe_pos_search" else key): v for key, v in c["cfg"].items()}
    rng = np.random.default_rng(100 + case)
    for it in range(1, 5):
        groups = mls_groups(P["n"], c.get("group", 4), rng)  # e.g. six groups of 4 and one of 1
        o = _oracle_from_engine(e, orc, amp)
        o.iteration(orc.mls_params(P["box"], iter=it, **okw), groups)
        ferr = e.mls_iteration(it, groups)
        g = e.mls_get_state()
        tag = f"case {case} iter {it}"
        assert np.array_equal(g["roi"], o.roi), tag
        assert rel_err(e.get_object(), o.obj) < 1e-5, tag
        assert rel_err(e.get_probe(), o.probe) < 1e-5, tag
        assert rel_err(g["variable_mode"], o.vmode) < 1e-4, tag
        # vectors compared by norm: single evolution/step entries are differences of O(1) terms
        assert rel_err(g["evolution"][0], o.evo[0]) < 1e-5, tag
        assert rel_err(g["evolution"][1], o.evo[1]) < 1e-4, tag
        assert np.max(np.abs(g["positions"] - o.pos)) < 1e-4, tag
        assert rel_err(g["beta_object"], o.beta_obj) < 1e-4, tag
        assert rel_err(g["beta_probe"], o.beta_probe) < 1e-4, tag
        if it < 10:
            assert rel_err(ferr, o.ferr) < 1e-5, tag


def test_mls_deterministic_and_converging(gpu):
    P = _problem(k=2, seed=7)
    runs = []
    for _ in range(2):
        e = _engine_for(P, P["state"].amp, probe_pos_search=5)
        rng = np.random.default_rng(3)
        errs = []
        for it in range(1, 10):
            errs.append(float(np.mean(e.mls_iteration(it, mls_groups(P["n"], 4, rng)))))
        runs.append((e.get_object(), e.get_probe(), errs))
        e.close()
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    errs = runs[0][2]
    assert errs[-1] < errs[0]


@pytest.mark.parametrize("m,k", [(128, 1), (128, 3), (256, 3)])
def test_mls_forked_schedule_is_bitwise_repeatable(gpu, m, k):
    """At 128^2 and 256^2 the position step, the variable-mode update, the evolution dots and the
    next group's illumination probes run on a second stream beside the main one (position search
    from iteration 2, variable probe and intensity from iteration 3): two runs in one process
    agree bit for bit in every state array."""
    P = _problem(k=k, m=m, n_side=4, seed=11)
    runs = []
    for _ in range(2):
        e = _engine_for(P, P["state"].amp, probe_pos_search=2, variable_probe=1, variable_intensity=1)
        rng = np.random.default_rng(5)
        for it in range(1, 6):
            e.mls_iteration(it, mls_groups(P["n"], 4, rng))
        g = e.mls_get_state()
        runs.append((e.get_object(), e.get_probe(), g["variable_mode"], g["positions"], g["evolution"],
                     g["beta_object"], g["beta_probe"]))
        e.close()
    for name, a, b in zip(("object", "probe", "variable mode", "positions", "evolution", "beta_object",
                           "beta_probe"), *runs):
        assert np.array_equal(a, b), name


def test_mls_rejects_bad_groups(gpu):
    from ptychopy_amd import _engine
    P = _problem(k=1)
    e = _engine_for(P, P["state"].amp)
    with pytest.raises(_engine.PtykError):
        e.mls_iteration(1, [np.array([0, 0])])
    with pytest.raises(_engine.PtykError):
        e.mls_iteration(1, [np.array([P["n"]])])
    with pytest.raises(_engine.PtykError):
        e.mls_iteration(0, [np.array([0])])


@pytest.mark.parametrize("m", [128, 256])
def test_mls_launch_shapes_give_the_same_iteration(gpu, orc, m):
    """The tuning knobs change launch shapes only: every modulus block width gives the same object
    and probe bit for bit (the per-column transforms and per-pixel modulus are the same arithmetic),
    every probe-mode chunk size the same within FP32 rounding of the chunk sums; each stays within
    1e-5 of the oracle."""
    from ptychopy_amd import _engine as E
    P = _problem(k=3, m=m, n_side=4, seed=21)
    rng0 = np.random.default_rng(8)
    groups = [mls_groups(P["n"], 9, rng0) for _ in range(3)]

    def run(cb, fpc):
        e = _engine_for(P, P["state"].amp, probe_pos_search=2)
        e.set_tuning(E.TUNE_MLS_MOD_CB, cb)
        e.set_tuning(E.TUNE_MLS_PUM_FPC, fpc)
        assert (e.get_tuning(E.TUNE_MLS_MOD_CB), e.get_tuning(E.TUNE_MLS_PUM_FPC)) == (cb, fpc)
        for it in (1, 2):
            e.mls_iteration(it, groups[it - 1])
        amp = e.get_frames().astype(np.float64)
        o = _oracle_from_engine(e, orc, amp)
        e.mls_iteration(3, groups[2])
        o.iteration(orc.mls_params(P["box"], iter=3, pps=2), groups[2])
        out = (e.get_object(), e.get_probe())
        assert rel_err(out[0], o.obj) < 1e-5 and rel_err(out[1], o.probe) < 1e-5
        e.close()
        return out

    base = run(16, 4)
    for cb in (8, 32):
        o, p = run(cb, 4)
        assert np.array_equal(o, base[0]) and np.array_equal(p, base[1]), cb
    for fpc in (2, 8):
        o, p = run(16, fpc)
        assert rel_err(o, base[0]) < 1e-6 and rel_err(p, base[1]) < 1e-6, fpc
    from ptychopy_amd._engine import PtykError
    e = _engine_for(P, P["state"].amp)
    for key, bad in ((E.TUNE_MLS_MOD_CB, 12), (E.TUNE_MLS_PUM_FPC, 1), (E.TUNE_MLS_ROW_WAVES, 2)):
        with pytest.raises(PtykError):
            e.set_tuning(key, bad)
    e.close()


@pytest.mark.parametrize("m,k", [(128, 1), (256, 1), (128, 3), (256, 3)])
def test_mls_row_waves_match_workgroup_tiles(gpu, orc, m, k):
    """The wave-tile row passes (one wave per tile of whole row groups, wave-local LDS exchanges)
    compute every row with the same arithmetic as the workgroup-tile passes: with one mode (no
    probe-mode update) five iterations through position refinement (the exit-wave pass also
    forms the patches' row derivatives), the variable probe and sub-pixel shifts give the same
    state bit for bit.  With more modes the probe-mode sums add their frames in another order:
    after six iterations (the rounding differences carried through every later group) within
    1e-5 of each other, and the sixth iteration within 1e-5 of the oracle either way."""
    from ptychopy_amd import _engine as E
    P = _problem(k=k, m=m, n_side=4, seed=23)
    rng0 = np.random.default_rng(9)
    groups = [mls_groups(P["n"], 5, rng0) for _ in range(6)]
    runs = []
    for waves in (1, 0):
        e = _engine_for(P, P["state"].amp, probe_pos_search=2, variable_probe=1, variable_intensity=1)
        e.set_tuning(E.TUNE_MLS_ROW_WAVES, waves)
        assert e.get_tuning(E.TUNE_MLS_ROW_WAVES) == waves
        for it in range(1, 6):
            e.mls_iteration(it, groups[it - 1])
        amp = e.get_frames().astype(np.float64)
        o = _oracle_from_engine(e, orc, amp)
        e.mls_iteration(6, groups[5])
        o.iteration(orc.mls_params(P["box"], iter=6, pps=2, variable_probe=1, variable_intensity=1), groups[5])
        assert rel_err(e.get_object(), o.obj) < 1e-5 and rel_err(e.get_probe(), o.probe) < 1e-5, waves
        g = e.mls_get_state()
        runs.append((e.get_object(), e.get_probe(), g["variable_mode"], g["positions"], g["evolution"]))
        e.close()
    names = ("object", "probe", "variable mode", "positions", "evolution")
    for name, a, b in zip(names, *runs):
        if k == 1:
            assert np.array_equal(a, b), name
        else:
            assert rel_err(a, b) < 1e-5, name


def test_mls_autotuner_persists_a_timing_table(gpu, tmp_path):
    import json
    from ptychopy_amd import autotune
    P = _problem(k=2, m=128, n_side=4, seed=22)
    path = tmp_path / "tune.json"
    best, times = autotune.tune_mls(P["pos"], 128, 2, P["shape"], P["box"], iters=1, path=path)
    assert ",".join(map(str, best)) in times and len(times) == 10
    table = json.loads(path.read_text())
    assert table[autotune.mls_key(P["n"], 128, 2, P["shape"], max(1, P["n"] // 100))] == times
    assert autotune.tune_mls(P["pos"], 128, 2, P["shape"], P["box"], path=path) == (best, times)


@pytest.mark.parametrize("m,k,group", [(128, 3, 4), (256, 5, 5), (64, 2, 4)])
def test_mls_group_graphs_equal_direct_launches(gpu, m, k, group):
    """Groups replayed from captured HIP graphs (one capture per group shape and flag set; groups
    of `group` frames and a smaller last one, so several shapes and replays occur in every
    iteration, with eager work and synchronisations between the iterations) give the same state
    bit for bit as the same iterations launched kernel by kernel, through position refinement,
    the variable probe and the Fourier-error iterations."""
    from ptychopy_amd import _engine as E
    P = _problem(k=k, m=m, n_side=4, seed=31)
    runs = []
    for graphs in (1, 0):
        e = _engine_for(P, P["state"].amp, probe_pos_search=2)
        e.set_tuning(E.TUNE_MLS_GRAPHS, graphs)
        rng = np.random.default_rng(6)
        ferrs = []
        for it in range(1, 7):
            ferrs.append(e.mls_iteration(it, mls_groups(P["n"], group, rng)))
            e.synchronize()
        g = e.mls_get_state()
        runs.append((e.get_object(), e.get_probe(), g["variable_mode"], g["positions"], g["evolution"],
                     np.array(ferrs)))
        e.close()
    for name, a, b in zip(("object", "probe", "variable mode", "positions", "evolution", "fourier errors"), *runs):
        assert np.array_equal(a, b), name
