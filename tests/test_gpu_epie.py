"""ePIE iteration on the GPU against the FP64 oracle (same state, same position order).

Single-mode 64^2 and 128^2 frames run the LDS-resident form (the exit wave held in LDS, the
canvas maximum tracked by location); 16^2 / 32^2 frames and the modal form run the staged form.
Both are checked, with and without texture-weight quantisation, with a second (unused) probe mode
whose intensities enter the probe maximum, and in the modal form."""
import numpy as np
import pytest

from conftest import rel_err
from problems import make_problem

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _engine_for(p, quant=0):
    from ptychopy_amd._engine import Engine
    e = Engine(p["n"], p["m"], p["shape"], n_modes=p["k"], tex_quantize=quant)
    e.set_frames(p["amp"])
    e.set_positions(p["pos"])
    e.clear_object()
    e.init_probe(beam_size=60e-9, dx_s=5e-9)
    return e


def _flags(i, rms=False):
    from ptychopy_amd import _engine as E
    f = E.FLAG_RMS if rms else 0
    if i <= 1:
        f |= E.FLAG_PHASE_CONSTRAINT
    if i >= 10:
        f |= E.FLAG_UPDATE_PROBE
    return f


def _check(e, orc, p, amp, order, flags, quant=0, k_use=1, modal=False):
    from ptychopy_amd import _engine as E
    o0, p0 = e.get_object(), e.get_probe()
    sc = e.get_scalars()
    st = orc.EpieState(o0, p0, sc[E.SCAL_OBJ_MAX], sc[E.SCAL_PROBE_MAX])
    err_g = e.epie_iteration(order, flags)
    err_o = orc.epie_iteration(st, amp, p["pos"], order, k_use=k_use, modal=modal,
                               phase_constraint=bool(flags & E.FLAG_PHASE_CONSTRAINT),
                               update_probe=bool(flags & E.FLAG_UPDATE_PROBE), rms=bool(flags & E.FLAG_RMS),
                               quantize=quant)
    errs = dict(obj=rel_err(e.get_object(), st.obj), probe=rel_err(e.get_probe(), st.probe))
    print("epie step errors", errs)
    assert errs["obj"] < TOL and errs["probe"] < TOL, errs
    sc1 = e.get_scalars()
    assert sc1[E.SCAL_OBJ_MAX] == pytest.approx(st.scal[0], rel=1e-5)
    assert sc1[E.SCAL_PROBE_MAX] == pytest.approx(st.scal[1], rel=1e-5)
    if flags & E.FLAG_RMS:
        assert err_g == pytest.approx(err_o, rel=1e-4)
    return errs


@pytest.mark.parametrize("m,n_side,seed,quant", [(32, 4, 0, 0), (64, 4, 1, 0), (128, 3, 2, 1), (16, 5, 3, 0),
                                                   (128, 4, 4, 0), (64, 3, 5, 1)])
def test_epie_iterations_match_oracle(gpu, orc, m, n_side, seed, quant):
    p = make_problem(m=m, n_side=n_side, seed=seed)
    rng = np.random.default_rng(seed)
    with _engine_for(p, quant) as e:
        amp = e.get_frames().astype(np.float64)
        for i in range(1, 12):
            order = rng.permutation(p["n"])
            if i in (1, 2, 10, 11):
                _check(e, orc, p, amp, order, _flags(i, rms=(i == 2)), quant)
            else:
                e.epie_iteration(order, _flags(i))


def test_epie_is_bitwise_deterministic(gpu):
    p = make_problem(m=64, n_side=3, seed=7)
    outs = []
    for _ in range(2):
        rng = np.random.default_rng(0)
        with _engine_for(p) as e:
            for i in range(1, 12):
                e.epie_iteration(rng.permutation(p["n"]), _flags(i))
            outs.append((e.get_object(), e.get_probe()))
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_epie_with_an_unused_second_mode_matches_oracle(gpu, orc):
    """Single-mode updates with K = 2 stored modes: the probe maximum also runs over mode 2."""
    from ptychopy_amd import _engine as E
    p = make_problem(m=128, n_side=3, seed=6, modes=2)
    rng = np.random.default_rng(6)
    with _engine_for(p) as e:
        pr = e.get_probe()
        pr[1] = 1.2 * pr[0].T  # brighter than mode 1 somewhere, so it sets the maximum at first
        e.set_probe(pr)
        e.refresh_maxima()
        amp = e.get_frames().astype(np.float64)
        for i in (1, 10, 11):
            _check(e, orc, p, amp, rng.permutation(p["n"]), _flags(i, rms=(i == 11)))


@pytest.mark.parametrize("m,seed", [(32, 8), (128, 9)])
def test_modal_epie_matches_oracle(gpu, orc, m, seed):
    """Modal ePIE (all modes updated against the summed intensity) after the modes are formed."""
    from ptychopy_amd import _engine as E
    p = make_problem(m=m, n_side=3, seed=seed, modes=2)
    rng = np.random.default_rng(seed)
    with _engine_for(p) as e:
        amp = e.get_frames().astype(np.float64)
        for i in range(1, 4):
            e.epie_iteration(rng.permutation(p["n"]), _flags(i))
        modal = E.FLAG_UPDATE_MODES | E.FLAG_UPDATE_PROBE
        e.epie_iteration(rng.permutation(p["n"]), modal)  # forms the modes (from the base probe)
        assert np.abs(e.get_probe()[1]).max() > 0
        for rms in (False, True):
            _check(e, orc, p, amp, rng.permutation(p["n"]), modal | (E.FLAG_RMS if rms else 0), k_use=2, modal=True)
