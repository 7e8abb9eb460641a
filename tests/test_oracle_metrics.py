"""Pins of the oracle's normalised object error (oracle.cpp orc_object_error) by closed forms."""
import numpy as np


def _setup(seed=0, m=16, rows=40, cols=44):
    rng = np.random.default_rng(seed)
    truth = rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))
    probe = np.exp(-((np.mgrid[0:m, 0:m] - m / 2) ** 2).sum(0) / 30.0)[None] * (1 + 0.5j)
    roi = np.array([[3, 4], [10, 12], [20, 25], [5, 22]], np.int32)
    cov = np.zeros((rows, cols))
    for r0, c0 in roi:
        cov[r0:r0 + m, c0:c0 + m] += np.abs(probe[0]) ** 2
    return rng, truth, probe, roi, cov > 0.01 * cov.max()


def test_global_scale_and_phase_are_free(orc):
    _, truth, probe, roi, mask = _setup()
    g0 = 0.7 * np.exp(1.1j)
    e, g, cnt = orc.object_error(truth / g0, truth, probe, roi)
    assert e < 1e-14 and abs(g - g0) < 1e-12 and cnt == mask.sum()


def test_orthogonal_perturbation_closed_form(orc):
    rng, truth, probe, roi, mask = _setup(1)
    d = rng.standard_normal(truth.shape) + 1j * rng.standard_normal(truth.shape)
    t, dm = truth[mask], d[mask]
    dm = dm - np.vdot(t, dm) / np.vdot(t, t) * t  # <T, D> = 0 on the region
    d[mask] = dm
    e, g, _ = orc.object_error(truth + d, truth, probe, roi)
    nt, nd = np.sum(np.abs(t) ** 2), np.sum(np.abs(dm) ** 2)
    assert np.isclose(g, nt / (nt + nd), rtol=1e-12)
    assert np.isclose(e, np.sqrt(nd / (nt + nd)), rtol=1e-12)


def test_pixels_outside_the_illuminated_region_do_not_count(orc):
    rng, truth, probe, roi, mask = _setup(2)
    o = truth * (1.0 + 0.01 * rng.standard_normal(truth.shape))
    e1 = orc.object_error(o, truth, probe, roi)[0]
    o2 = o.copy()
    o2[~mask] = 1e3
    assert orc.object_error(o2, truth, probe, roi)[0] == e1
