"""The engine's normalised object error against the FP64 oracle (through the C ABI)."""
import numpy as np
import pytest

from problems import make_mls_problem, make_problem

pytestmark = pytest.mark.gpu


def test_object_error_matches_oracle(gpu, orc):
    from ptychopy_amd._engine import Engine
    P = make_problem(m=32, n_side=4, seed=3)
    rng = np.random.default_rng(0)
    truth = P["obj"]
    est = (truth * 0.8 * np.exp(0.4j) + 0.05 * (rng.standard_normal(truth.shape)
                                               + 1j * rng.standard_normal(truth.shape))).astype(np.complex64)
    probe = np.stack([P["probe"], 0.3 * P["probe"][::-1]])
    with Engine(P["n"], 32, P["shape"], n_modes=2) as e:
        e.set_positions(P["pos"])
        e.set_object(est)
        e.set_probe(probe)
        e.set_true_object(truth)
        got = e.object_error(0.01)
        ref = orc.object_error(est, truth, probe.astype(np.complex64), e.get_roi(), 0.01)
    assert got[2] == ref[2]
    assert abs(got[0] - ref[0]) <= 1e-6 * ref[0]
    assert abs(got[1] - ref[1]) <= 1e-6 * abs(ref[1])


def test_object_error_on_the_mls_rois(gpu, orc):
    from ptychopy_amd import _engine
    P = make_mls_problem(m=32, n_side=4, k=2, seed=1, photons=1e5, zero_evo01=False)
    st = P["state"]
    with _engine.Engine(P["n"], 32, P["shape"], n_modes=2) as e:
        e.set_frames(st.amp)
        e.set_object(P["obj0"])
        e.set_probe(P["probe0"])
        e.mls_begin(_engine.mls_config(P["box"]), P["pos"], P["vmode0"], P["evo1_0"], P["frame_power"])
        e.set_true_object(P["obj0"])
        got = e.object_error(0.01)
        g = e.mls_get_state()
        ref = orc.object_error(e.get_object(), P["obj0"].astype(np.complex64), e.get_probe(), g["roi"], 0.01)
    # initEvo only rescales the object: the error is ~0 and gamma undoes the scale
    assert got[0] < 1e-6 and got[2] == ref[2]
    assert abs(got[1] - ref[1]) <= 1e-6 * abs(ref[1])
