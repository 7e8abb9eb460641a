"""Whole reconstructions on the GPU and in the FP64 oracle, run independently from the same start
(same data, same position order and groups): the final normalised object error agrees within
1 % (relative), the bar for the final error metric; the engine's own metric is compared with
the oracle's on the oracle's result."""
import numpy as np
import pytest

from problems import make_mls_problem, make_problem, mls_groups

pytestmark = pytest.mark.gpu

FINAL_TOL = 0.01


def _flags(i):
    from ptychopy_amd import _engine as E
    f = 0
    if i <= 1:
        f |= E.FLAG_PHASE_CONSTRAINT
    if i >= 10:
        f |= E.FLAG_UPDATE_PROBE
    return f


def _dm_epie_engine(p):
    from ptychopy_amd._engine import Engine
    e = Engine(p["n"], p["m"], p["shape"])
    e.set_frames(p["amp"])
    e.set_positions(p["pos"])
    e.clear_object()
    e.set_probe(p["probe"][None])  # the true probe: the runs converge, so their errors are informative
    e.refresh_maxima()
    e.set_true_object(p["obj"])
    return e


def test_dm_final_object_error_matches_oracle(gpu, orc):
    from ptychopy_amd import _engine as E
    p = make_problem(m=32, n_side=5, seed=21, jitter=1.0)
    with _dm_epie_engine(p) as e:
        amp = e.get_frames().astype(np.float64)
        o, pr = e.get_object().astype(complex), e.get_probe()[0].astype(complex)
        sc = e.get_scalars()
        fac = np.array([np.inf, 1 / np.float32(sc[E.SCAL_PROBE_MAX])])
        psi = np.zeros((p["n"], 32, 32))
        for i in range(1, 41):
            e.dm_iteration(_flags(i))
            o, pr, psi, fac, _ = orc.dm_iteration(o, pr, psi, amp, p["pos"], fac, phase_constraint=i <= 1,
                                                  update_probe=i >= 10)
        eg = e.object_error(0.01)[0]
        roi = e.get_roi()
    eo = orc.object_error(o, p["obj"], pr[None], roi, 0.01)[0]
    print("DM final e", eg, eo)
    assert eo < 0.8  # the run reconstructs something (a zero object has e = 1)
    assert abs(eg - eo) <= FINAL_TOL * eo


def test_epie_final_object_error_matches_oracle(gpu, orc):
    from ptychopy_amd import _engine as E
    p = make_problem(m=32, n_side=5, seed=22, jitter=1.0)
    rng = np.random.default_rng(22)
    with _dm_epie_engine(p) as e:
        amp = e.get_frames().astype(np.float64)
        sc = e.get_scalars()
        st = orc.EpieState(e.get_object(), e.get_probe(), sc[E.SCAL_OBJ_MAX], sc[E.SCAL_PROBE_MAX])
        for i in range(1, 31):
            order = rng.permutation(p["n"])
            f = _flags(i)
            e.epie_iteration(order, f)
            orc.epie_iteration(st, amp, p["pos"], order, phase_constraint=bool(f & E.FLAG_PHASE_CONSTRAINT),
                               update_probe=bool(f & E.FLAG_UPDATE_PROBE))
        eg = e.object_error(0.01)[0]
        roi = e.get_roi()
    eo = orc.object_error(st.obj, p["obj"], st.probe, roi, 0.01)[0]
    print("ePIE final e", eg, eo)
    assert eo < 0.8
    assert abs(eg - eo) <= FINAL_TOL * eo


def test_mls_final_object_error_matches_oracle(gpu, orc):
    from ptychopy_amd import _engine
    P = make_mls_problem(m=32, n_side=5, k=2, seed=23, phase_object=False, photons=2e5, zero_evo01=False)
    truth = P["obj0"]
    start = np.ones(P["shape"], np.complex128)
    rng = np.random.default_rng(23)
    k, m, _ = P["probe0"].shape
    with _engine.Engine(P["n"], m, P["shape"], n_modes=k) as e:
        e.set_frames(P["state"].amp)
        amp = e.get_frames().astype(np.float64)
        e.set_object(start)
        e.set_probe(P["probe0"])
        e.set_true_object(truth)
        e.mls_begin(_engine.mls_config(P["box"], probe_pos_search=99), P["pos"], P["vmode0"], P["evo1_0"],
                    P["frame_power"])
        o = orc.MlsState(start, P["probe0"], P["vmode0"], P["evo1_0"], P["pos"], amp)
        o.init(P["frame_power"])
        for it in range(1, 13):
            groups = mls_groups(P["n"], 4, rng)
            e.mls_iteration(it, groups)
            o.iteration(orc.mls_params(P["box"], iter=it, pps=99), groups)
        eg = e.object_error(0.01)[0]
        roi = e.mls_get_state()["roi"]
    eo = orc.object_error(o.obj, truth, o.probe, roi, 0.01)[0]
    print("MLs final e", eg, eo)
    assert eo < 0.9
    assert abs(eg - eo) <= FINAL_TOL * eo
