"""Warm-start plumbing on the GPU: on-device input preprocessing and patch stitching, against
numpy restatements (the stitch is a weighted average, so float32 sums agree to ~1e-6)."""
import numpy as np
import pytest

from ptychopy_amd import _engine, warmstart, sim

pytestmark = pytest.mark.gpu


def stitch_reference(pred, roi, weight, canvas, thr, fill):
    m = pred.shape[-1]
    num = np.zeros(canvas, np.complex128)
    den = np.zeros(canvas, np.float64)
    for j, (r0, c0) in enumerate(roi):
        num[r0:r0 + m, c0:c0 + m] += weight * pred[j, 0] * np.exp(1j * pred[j, 1].astype(np.float64))
        den[r0:r0 + m, c0:c0 + m] += weight
    out = np.full(canvas, fill, np.complex128)
    ok = den > thr * weight.max()
    out[ok] = num[ok] / den[ok]
    return out, ok


def test_stitch_matches_weighted_average():
    rng = np.random.default_rng(0)
    m, k = 32, 2
    pos = sim.raster_positions(6, 9.0, rng, jitter=1.0, origin=(2.0, 3.0))
    canvas = (90, 100)  # wider than the scan: some pixels keep the fill value
    n = len(pos)
    with _engine.Engine(n, m, canvas, n_modes=k) as e:
        e.set_positions(pos)
        probe = (rng.standard_normal((k, m, m)) + 1j * rng.standard_normal((k, m, m))).astype(np.complex64)
        e.set_probe(probe)
        roi = e.get_roi()
        sel = np.array([0, 3, 7, 8, 20, 35, 5], np.int32)
        pred = np.stack([rng.uniform(0.5, 1.5, (len(sel), m, m)), rng.uniform(-3, 3, (len(sel), m, m))],
                        axis=1).astype(np.float32)
        e.stitch_patches(pred, sel, rel_threshold=0.05, fill=0.25 - 0.5j)
        got = e.get_object()
    w = (np.abs(probe.astype(np.complex128)) ** 2).sum(axis=0)
    ref, ok = stitch_reference(pred, roi[sel], w, canvas, 0.05, 0.25 - 0.5j)
    assert (~ok).any() and ok.any()
    np.testing.assert_array_equal(got[~ok], np.complex64(0.25 - 0.5j))
    err = np.abs(got[ok] - ref[ok]).max() / np.abs(ref[ok]).max()
    assert err <= 1e-5, err


def test_stitch_rejects_bad_frames():
    with _engine.Engine(4, 16, (40, 40)) as e:
        e.set_positions(np.array([[0, 0], [0, 8], [8, 0], [8, 8]], np.float32))
        with pytest.raises(_engine.PtykError, match="out of range"):
            e.stitch_patches(np.zeros((1, 2, 16, 16), np.float32), np.array([4], np.int32))


def test_predict_amplitudes_equals_host_preprocessing():
    rng = np.random.default_rng(1)
    s = 32
    amp = np.sqrt(rng.poisson(20.0 * rng.uniform(0, 1, (5, s, s)) ** 4)).astype(np.float32)
    amp[2] = 0.0  # an all-dark frame maps to a zero input
    with warmstart.Net(size=s, c=4, batch=2, seed=2) as net:
        a = net.predict_amplitudes(amp)
        b = net.predict(warmstart.preprocess(amp))
    assert np.linalg.norm(a - b) <= 1e-3 * np.linalg.norm(b)


def test_warm_start_places_patches_at_the_mls_rois():
    rng = np.random.default_rng(3)
    m, k, side = 32, 2, 4
    pos = sim.raster_positions(side, 6.0, rng, jitter=0.0).astype(np.float32)
    canvas = (70, 70)
    n = len(pos)
    amp = rng.uniform(0, 5, (n, m, m)).astype(np.float32)
    probe = sim.hermite_gaussian_modes(m, k, 8.0)
    with _engine.Engine(n, m, canvas, n_modes=k) as e, warmstart.Net(size=m, c=4, batch=4, seed=4) as net:
        e.set_frames(amp)
        e.set_object(np.ones(canvas, np.complex64))
        e.set_probe(probe)
        box = (1, canvas[1], 1, canvas[0])
        vmode = (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))).astype(np.complex64)
        e.mls_begin(_engine.mls_config(box), pos, vmode, 1e-6 * rng.standard_normal(n),
                    (amp.astype(np.float64) ** 2).sum(axis=(1, 2)))
        roi = e.mls_get_state()["roi"]
        probe_now = e.get_probe()
        sel = warmstart.lattice_frames(side, side, 2)
        pred = warmstart.warm_start(e, net, amp, sel)
        got = e.get_object()
    w = (np.abs(probe_now.astype(np.complex128)) ** 2).sum(axis=0)
    ref, ok = stitch_reference(pred, roi[sel], w, canvas, 1e-3, 1.0)
    err = np.abs(got[ok] - ref[ok]).max() / np.abs(ref[ok]).max()
    assert err <= 1e-5, err
    np.testing.assert_array_equal(got[~ok], np.complex64(1.0))
