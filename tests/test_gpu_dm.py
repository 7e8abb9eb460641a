"""DM iteration on the GPU against the FP64 oracle, starting each check from the GPU's state.

Tolerance: per-iteration object and probe within 1e-5 relative (L2 norm) of
the oracle's FP64 result from identical inputs (dequantised FP16 amplitudes, the same
positions and the same cached factors).  The exit-wave state psi_new = psi + Pf - 2 P O_r
is a difference of terms of the size of the transform input x = 2 P O_r - psi, so in any
FP32 evaluation (the reference's included) its error scales with ||x||: it is held to 1e-5
of max(||psi_new||, ||x||).
"""
import numpy as np
import pytest

from conftest import rel_err
from problems import make_problem

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _engine_for(p, beam=None):
    from ptychopy_amd._engine import Engine
    e = Engine(p["n"], p["m"], p["shape"], n_modes=p["k"])
    e.set_frames(p["amp"])
    e.set_positions(p["pos"])
    e.clear_object()
    e.init_probe(beam_size=beam or 60e-9, dx_s=5e-9)
    return e


def _flags(i, phase_constraint=1, update_probe=10):
    from ptychopy_amd import _engine as E
    f = 0
    if i <= phase_constraint:
        f |= E.FLAG_PHASE_CONSTRAINT
    if i >= update_probe:
        f |= E.FLAG_UPDATE_PROBE
    return f


def _factors_for_next_call(e, first_call):
    """Cached DM factors the next iteration uses: on the first call, {1/objMax, 1/probeMax} from
    the current maxima (DM.cu:88-92, objMax = 0 after clearing gives inf); later the cache."""
    from ptychopy_amd import _engine as E
    sc = e.get_scalars()
    if first_call:
        om = np.float32(sc[E.SCAL_OBJ_MAX])
        return np.array([np.inf if om == 0 else 1 / om, 1 / np.float32(sc[E.SCAL_PROBE_MAX])])
    return np.array([sc[E.SCAL_DM_FACTOR_OBJ], sc[E.SCAL_DM_FACTOR_PROBE]])


def _check_step(e, orc, p, amp, flags, psi0, first_call=False):
    from ptychopy_amd import _engine as E
    o0 = e.get_object()
    p0 = e.get_probe()[0]
    fac = _factors_for_next_call(e, first_call)
    e.dm_iteration(flags)
    o1, p1, s1 = e.get_object(), e.get_probe()[0], e.get_psi()
    sc1 = e.get_scalars()
    ro, rp, rs, rf, rsum = orc.dm_iteration(o0, p0, psi0, amp, p["pos"], fac,
                                            phase_constraint=bool(flags & E.FLAG_PHASE_CONSTRAINT),
                                            update_probe=bool(flags & E.FLAG_UPDATE_PROBE))
    m = p["m"]
    x = np.stack([2 * p0 * o0[r:r + m, c:c + m] for r, c in e.get_roi()]).astype(complex) - psi0
    scale = max(np.linalg.norm(rs.ravel()), np.linalg.norm(x.ravel()))
    errs = dict(obj=rel_err(o1, ro), probe=rel_err(p1, rp),
                psi=float(np.linalg.norm((s1 - rs).ravel()) / scale),
                psi_rel=rel_err(s1, rs), obj_upd=rel_err(o1.astype(complex) - o0, ro - o0))
    print("dm step errors", errs)
    assert errs["obj"] < TOL and errs["probe"] < TOL and errs["psi"] < TOL, errs
    assert sc1[E.SCAL_OBJ_MAX] == pytest.approx(rsum[0], rel=1e-5)
    assert sc1[E.SCAL_PROBE_MAX] == pytest.approx(rsum[1], rel=1e-5)
    g = np.array([sc1[E.SCAL_DM_FACTOR_OBJ], sc1[E.SCAL_DM_FACTOR_PROBE]])
    assert np.allclose(g, rf, rtol=1e-5)
    return s1, errs


@pytest.mark.parametrize("m,n_side,seed", [(32, 4, 0), (64, 5, 1), (128, 3, 2), (16, 6, 3)])
def test_dm_first_iterations_match_oracle(gpu, orc, m, n_side, seed):
    p = make_problem(m=m, n_side=n_side, seed=seed)
    with _engine_for(p) as e:
        amp = e.get_frames().astype(np.float64)
        # DM.cu:88-92: first call caches {1/objMax, 1/probeMax} = {1/0, 1/max|P|^2}
        sc = e.get_scalars()
        from ptychopy_amd import _engine as E
        fac0 = np.array([1 / np.float32(sc[E.SCAL_OBJ_MAX]) if sc[E.SCAL_OBJ_MAX] else np.inf,
                         1 / np.float32(sc[E.SCAL_PROBE_MAX])])
        e.set_psi(np.zeros((p["n"], m, m), np.complex64))
        e.set_scalars(np.array([sc[0], sc[1], 0, 0, fac0[0], fac0[1], 0, 0]))
        psi = np.zeros((p["n"], m, m))
        for i in (1, 2, 3):
            psi, errs = _check_step(e, orc, p, amp, _flags(i), psi)


def test_dm_probe_update_iterations_match_oracle(gpu, orc):
    p = make_problem(m=64, n_side=5, seed=4)
    with _engine_for(p) as e:
        amp = e.get_frames().astype(np.float64)
        for i in range(1, 10):
            e.dm_iteration(_flags(i))
        psi = e.get_psi()
        for i in (10, 11):
            psi, errs = _check_step(e, orc, p, amp, _flags(i), psi)


def test_dm_is_bitwise_deterministic(gpu):
    p = make_problem(m=64, n_side=4, seed=5)
    outs = []
    for _ in range(2):
        with _engine_for(p) as e:
            for i in range(1, 13):
                e.dm_iteration(_flags(i))
            outs.append((e.get_object(), e.get_probe(), e.get_psi()))
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_dm_many_iterations_stay_close_to_oracle(gpu, orc):
    # 20 iterations run independently on both sides from the same start: drift stays small.
    from ptychopy_amd import _engine as E
    p = make_problem(m=32, n_side=5, seed=6)
    with _engine_for(p) as e:
        amp = e.get_frames().astype(np.float64)
        o, pr = e.get_object().astype(complex), e.get_probe()[0].astype(complex)
        sc = e.get_scalars()
        fac = np.array([np.inf, 1 / np.float32(sc[E.SCAL_PROBE_MAX])])
        psi = np.zeros((p["n"], 32, 32))
        for i in range(1, 21):
            e.dm_iteration(_flags(i))
            o, pr, psi, fac, _ = orc.dm_iteration(o, pr, psi, amp, p["pos"], fac,
                                                  phase_constraint=i <= 1, update_probe=i >= 10)
        assert rel_err(e.get_object(), o) < 1e-4
        assert rel_err(e.get_probe()[0], pr) < 1e-4


@pytest.mark.parametrize("m,n_side,wave", [(128, 4, 5), (256, 3, 2), (512, 2, 3), (128, 3, 0)])
def test_dm_wave_schedule_matches_oracle(gpu, orc, m, n_side, wave):
    """Wave-scheduled form (several waves per iteration) against the oracle."""
    from ptychopy_amd import _engine as E
    p = make_problem(m=m, n_side=n_side, seed=10 + m + wave)
    with _engine_for(p) as e:
        e.set_tuning(E.TUNE_DM_WAVE, wave)
        amp = e.get_frames().astype(np.float64)
        psi = np.zeros((p["n"], m, m))
        for i in range(1, 4):
            psi, errs = _check_step(e, orc, p, amp, _flags(i), psi, first_call=(i == 1))
        assert e.get_tuning(E.TUNE_DM_WAVE) == (wave or e.get_tuning(E.TUNE_DM_WAVE))
        for i in range(4, 10):
            e.dm_iteration(_flags(i))
        psi = e.get_psi()
        for i in (10, 11):
            psi, errs = _check_step(e, orc, p, amp, _flags(i), psi)


def test_dm_wave_and_single_pass_agree(gpu):
    from ptychopy_amd import _engine as E
    p = make_problem(m=128, n_side=4, seed=21)
    res = []
    for wave in (-1, 3, 16):
        with _engine_for(p) as e:
            e.set_tuning(E.TUNE_DM_WAVE, wave)
            for i in range(1, 13):
                e.dm_iteration(_flags(i))
            res.append((e.get_object(), e.get_probe()))
    for o, pr in res[1:]:
        assert rel_err(o, res[0][0]) < 1e-5
        assert rel_err(pr, res[0][1]) < 1e-5


def test_dm_wave_is_bitwise_deterministic(gpu):
    from ptychopy_amd import _engine as E
    p = make_problem(m=256, n_side=3, seed=22)
    outs = []
    for _ in range(2):
        with _engine_for(p) as e:
            e.set_tuning(E.TUNE_DM_WAVE, 4)
            for i in range(1, 12):
                e.dm_iteration(_flags(i))
            outs.append((e.get_object(), e.get_probe(), e.get_psi()))
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_dm_timed_iteration_reports_every_pass(gpu):
    from ptychopy_amd import _engine as E
    p = make_problem(m=128, n_side=4, seed=23)
    with _engine_for(p) as e:
        ms = e.dm_iteration_timed(_flags(12))
    assert ms.shape == (len(E.DM_KERNELS),)
    assert np.all(ms[[0, 1, 2, 3, 4]] > 0)
