"""Parity at the benchmark workloads themselves (full sizes), through the C ABI, against the FP64
oracle started from the GPU's own state.

  C0  ePIE 12 x 12 raster, 128^2 frames, 320^2 object: two iterations over all 144 positions
  C1  DM 30 x 30 raster, 512^2 frames, 1027^2 canvas, 64-frame waves: one iteration, probe update on
  C2  MLs 51 x 51 raster, 256^2 frames, 5 modes, 101 groups: one iteration with position refinement,
      variable probe and intensity
  C3  16 DM scans of uneven size: batched (captured graphs) bitwise equal to standalone runs for
      three iterations; the largest 512^2 scan one iteration against the oracle

Tolerances as in the smaller parity tests: object and probe within 1e-5 relative (L2) per
iteration; DM exit waves within 1e-5 of max(||psi_new||, ||x||); ROI origins bit-exact.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _say(msg):
    print(msg, flush=True)


def _dm_compare(e, orc, pos, amp, flags, first_call=False):
    """One DM iteration on the GPU and in the oracle from the GPU's state; returns the errors."""
    from ptychopy_amd import _engine as E
    o0, p0, psi0 = e.get_object(), e.get_probe()[0], e.get_psi()
    sc = e.get_scalars()
    if first_call:
        om = np.float32(sc[E.SCAL_OBJ_MAX])
        fac = np.array([np.inf if om == 0 else 1 / om, 1 / np.float32(sc[E.SCAL_PROBE_MAX])])
    else:
        fac = np.array([sc[E.SCAL_DM_FACTOR_OBJ], sc[E.SCAL_DM_FACTOR_PROBE]])
    e.dm_iteration(flags)
    o1, p1, s1 = e.get_object(), e.get_probe()[0], e.get_psi()
    ro, rp, rs, _, _ = orc.dm_iteration(o0, p0, psi0, amp, pos, fac,
                                        phase_constraint=bool(flags & E.FLAG_PHASE_CONSTRAINT),
                                        update_probe=bool(flags & E.FLAG_UPDATE_PROBE))
    m = p0.shape[0]
    roi = e.get_roi()
    # ||x|| with x = 2 P O_r - psi, frame by frame (no full-size temporary)
    xs = 0.0
    for j, (r, c) in enumerate(roi):
        xs += float(np.sum(np.abs(2 * p0.astype(np.complex128) * o0[r:r + m, c:c + m] - psi0[j]) ** 2))
    scale = max(np.linalg.norm(rs.ravel()), np.sqrt(xs))
    errs = dict(obj=rel_err(o1, ro), probe=rel_err(p1, rp), psi=float(np.linalg.norm((s1 - rs).ravel()) / scale))
    return errs


def test_c1_dm_iteration_matches_oracle(gpu, orc):
    from ptychopy_amd import workloads
    from ptychopy_amd import _engine as E
    _say("C1: generating the 30x30 / 512^2 DM workload")
    W = workloads.dm_c1()
    n, m = W["n"], W["m"]
    with E.Engine(n, m, W["canvas"]) as e:
        e.set_tuning(E.TUNE_DM_WAVE, 64)  # the benchmarked schedule: 15 waves of 64 frames
        e.set_frames(W["amp"])
        e.set_positions(W["pos"])
        e.clear_object()
        e.init_probe(W["beam"], W["dx_s"])
        amp = e.get_frames().astype(np.float64)
        for i in (1, 2, 3):
            e.dm_iteration(E.FLAG_PHASE_CONSTRAINT if i == 1 else 0)
        assert e.get_tuning(E.TUNE_DM_WAVE) == 64
        _say("C1: one iteration, probe update on, against the oracle")
        errs = _dm_compare(e, orc, W["pos"], amp, E.FLAG_UPDATE_PROBE)
    _say(f"C1 errors {errs}")
    assert errs["obj"] < TOL and errs["probe"] < TOL and errs["psi"] < TOL, errs


def test_c0_epie_iterations_match_oracle(gpu, orc):
    from ptychopy_amd import workloads
    from ptychopy_amd import _engine as E
    _say("C0: generating the 12x12 / 128^2 ePIE workload")
    W = workloads.epie_c0()
    n, m = W["n"], W["m"]
    rng = np.random.default_rng(20261200)
    with E.Engine(n, m, W["canvas"]) as e:
        e.set_frames(W["amp"])
        e.set_positions(W["pos"])
        e.clear_object()
        e.init_probe(W["beam"], W["dx_s"])
        amp = e.get_frames().astype(np.float64)
        sc = e.get_scalars()
        st = orc.EpieState(e.get_object(), e.get_probe(), sc[E.SCAL_OBJ_MAX], sc[E.SCAL_PROBE_MAX])
        for i, flags in ((1, E.FLAG_PHASE_CONSTRAINT), (2, E.FLAG_UPDATE_PROBE | E.FLAG_RMS)):
            order = rng.permutation(n)
            _say(f"C0: iteration {i}, all {n} positions")
            err_g = e.epie_iteration(order, flags)
            err_o = orc.epie_iteration(st, amp, W["pos"], order, phase_constraint=bool(flags & E.FLAG_PHASE_CONSTRAINT),
                                       update_probe=bool(flags & E.FLAG_UPDATE_PROBE), rms=bool(flags & E.FLAG_RMS))
            errs = dict(obj=rel_err(e.get_object(), st.obj), probe=rel_err(e.get_probe(), st.probe))
            _say(f"C0 iteration {i} errors {errs}")
            assert errs["obj"] < TOL and errs["probe"] < TOL, errs
            if flags & E.FLAG_RMS:
                assert err_g == pytest.approx(err_o, rel=1e-4)
            # the next iteration starts both from the GPU's state
            sc = e.get_scalars()
            st = orc.EpieState(e.get_object(), e.get_probe(), sc[E.SCAL_OBJ_MAX], sc[E.SCAL_PROBE_MAX])


def test_c2_mls_iteration_matches_oracle(gpu, orc):
    from ptychopy_amd import workloads, _driver
    from ptychopy_amd import _engine as E
    _say("C2: generating the 51x51 / 256^2 / 5-mode MLs workload")
    W = workloads.mls_c2()
    n, m, k = W["n"], W["m"], W["k"]
    rows, cols = W["canvas"]
    rng = np.random.default_rng(20261204)
    with E.Engine(n, m, (rows, cols), n_modes=k) as e:
        e.set_frames(W["amp"])
        amp = e.get_frames().astype(np.float64)
        e.set_object(np.ones((rows, cols), np.complex64))
        e.set_probe(W["probe_true"])
        vmode = (rng.uniform(0, 1, (m, m)) * np.exp(2j * np.pi * rng.uniform(0, 1, (m, m)))).astype(np.complex64)
        evo1 = 1e-6 * rng.standard_normal(n)
        cfg = E.mls_config(W["box"])
        e.mls_begin(cfg, W["pos"], vmode, evo1, (amp ** 2).sum(axis=(1, 2)))
        parts = _driver.mls_partitions(n, 100, 1)
        assert len(parts[0]) == 101

        def groups(i):
            p = parts[i % len(parts)]
            return [p[j] for j in rng.permutation(len(p))]

        for it in (1, 2):
            e.mls_iteration(it, groups(it))
        # the compared iteration refines positions (iteration >= probe_pos_search)
        it = int(cfg.probe_pos_search)
        gr = groups(it)
        g = e.mls_get_state()
        st = orc.MlsState(e.get_object(), e.get_probe(), g["variable_mode"], g["evolution"][1],
                          g["positions_start"], amp)
        st.pos = g["positions"].copy()
        st.evo = g["evolution"].astype(np.float64).copy()
        st.illum = g["illumination"].astype(np.float64)
        st.max_illum = float(g["max_illumination"])
        st.roi = g["roi"].copy()
        st.sub = g["subpixel"].copy()
        st.beta_obj = g["beta_object"].astype(np.complex128)
        st.beta_probe = g["beta_probe"].astype(np.complex128)
        _say(f"C2: iteration {it} on the GPU and in the oracle (101 groups)")
        e.mls_iteration(it, gr)
        st.iteration(orc.mls_params(W["box"], iter=it), gr)
        g1 = e.mls_get_state()
        errs = dict(obj=rel_err(e.get_object(), st.obj), probe=rel_err(e.get_probe(), st.probe),
                    vmode=rel_err(g1["variable_mode"], st.vmode),
                    pos=float(np.abs(g1["positions"] - st.pos).max()))
    _say(f"C2 errors {errs}")
    assert np.array_equal(g1["roi"], st.roi)
    assert errs["obj"] < TOL and errs["probe"] < TOL, errs
    assert errs["vmode"] < 1e-4 and errs["pos"] < 1e-4, errs


def test_c3_batch_equals_standalone_and_largest_scan_matches_oracle(gpu, orc):
    from ptychopy_amd import workloads, autotune
    from ptychopy_amd import _engine as E
    _say("C3: generating the 16 scans")
    scans = workloads.batch_c3()
    waves = [autotune.tune_dm_wave(s["pos"], s["m"], s["canvas"], iters=1)[0] for s in scans]

    def make(s, w):
        e = E.Engine(s["n"], s["m"], s["canvas"])
        e.set_frames(s["amp"])
        e.set_positions(s["pos"])
        e.clear_object()
        e.init_probe(s["m"] * 24.0 / 128.0 * 1e-8, 1e-8)
        e.set_tuning(E.TUNE_DM_WAVE, w)
        return e

    def flags(i):
        return E.FLAG_PHASE_CONSTRAINT if i <= 1 else E.FLAG_UPDATE_PROBE

    _say("C3: standalone runs")
    alone = []
    for s, w in zip(scans, waves):
        e = make(s, w)
        for i in (1, 2, 3):
            e.dm_iteration(flags(i))
        alone.append((e.get_object(), e.get_probe(), e.get_psi()))
        e.close()
    _say("C3: batched runs")
    es = [make(s, w) for s, w in zip(scans, waves)]
    with E.Batch(es) as b:
        for i in (1, 2, 3):
            b.dm_iteration(flags(i))
        b.synchronize()
        assert b.graph_count >= 1
        for j, (e, (o, p, psi)) in enumerate(zip(es, alone)):
            assert np.array_equal(e.get_object(), o), f"scan {j} object"
            assert np.array_equal(e.get_probe(), p), f"scan {j} probe"
            assert np.array_equal(e.get_psi(), psi), f"scan {j} psi"
    del alone
    for e in es:
        e.close()
    big = max(range(len(scans)), key=lambda j: (scans[j]["m"], scans[j]["n"]))
    s = scans[big]
    _say(f"C3: scan {big} ({s['n']} frames of {s['m']}^2) one iteration against the oracle")
    with make(s, waves[big]) as e:
        amp = e.get_frames().astype(np.float64)
        e.dm_iteration(flags(1))
        e.dm_iteration(flags(2))
        errs = _dm_compare(e, orc, s["pos"], amp, flags(3))
    _say(f"C3 errors {errs}")
    assert errs["obj"] < TOL and errs["probe"] < TOL and errs["psi"] < TOL, errs
