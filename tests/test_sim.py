"""Statistics of the seeded simulator (ptychopy_amd/sim.py): the random-field object, the probes
and the Poisson sampling, each against the closed form of its construction; Parseval for the
simulated intensities."""
import numpy as np
import pytest

from ptychopy_amd import sim


def test_seeded_streams_are_reproducible():
    a = sim.random_field_object((64, 64), sim.rng_for(1, 0))
    b = sim.random_field_object((64, 64), sim.rng_for(1, 0))
    c = sim.random_field_object((64, 64), sim.rng_for(2, 0))
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)


@pytest.mark.parametrize("phase_std,corr", [(0.5, 6.0), (0.2, 3.0), (1.0, 12.0)])
def test_random_field_phase_statistics(phase_std, corr):
    """Phase: white noise through a Gaussian filter of sigma = corr, rescaled to the requested std;
    its autocorrelation is exp(-d^2 / (4 corr^2)) (the filter's kernel convolved with itself)."""
    o = sim.random_field_object((512, 512), np.random.default_rng(0), phase_std=phase_std, phase_corr=corr)
    ph = np.angle(o.astype(np.complex128))
    # angle() wraps the few samples beyond +-pi (0.2% at std 1), which moves mean and std slightly
    assert abs(ph.std() - phase_std) <= 0.01 * phase_std
    assert abs(ph.mean()) <= 5e-3
    f = np.fft.fft2(ph - ph.mean())
    ac = np.real(np.fft.ifft2(np.abs(f) ** 2))
    ac /= ac[0, 0]
    # the sample autocorrelation's error grows with the number of px per correlation cell
    tol = 0.04 * max(1.0, corr / 6.0)
    for d in (corr / 2, corr, 2 * corr):
        di = int(round(d))
        rho = 0.5 * (ac[di, 0] + ac[0, di])
        assert abs(rho - np.exp(-di * di / (4 * corr * corr))) <= tol, (d, rho)


def test_random_field_amplitude_range():
    o = sim.random_field_object((256, 256), np.random.default_rng(1), amp_lo=0.8, amp_hi=1.0)
    a = np.abs(o)
    assert a.min() >= 0.8 - 1e-6 and a.max() <= 1.0 + 1e-6
    assert abs(a.mean() - 0.9) <= 0.01


def test_gaussian_probe_fwhm_and_defocus():
    m, fwhm = 128, 24.0
    p = sim.gaussian_probe(m, fwhm, defocus_edge=2.0).astype(np.complex128)
    row = np.abs(p[m // 2])
    # half-maximum crossings on either side of the centre, by linear interpolation
    half = row.max() / 2
    right = np.nonzero(row[m // 2:] < half)[0][0] + m // 2
    left = m // 2 - np.nonzero(row[:m // 2 + 1][::-1] < half)[0][0]
    xr = right - 1 + (row[right - 1] - half) / (row[right - 1] - row[right])
    xl = left + 1 - (row[left + 1] - half) / (row[left + 1] - row[left])
    assert abs((xr - xl) - fwhm) <= 0.05
    assert abs(row.max() - 1.0) <= 1e-6
    # phase 2 (r / (m/2))^2: zero at the centre, 2 rad at r = m/2 along an axis
    assert abs(np.angle(p[m // 2, m // 2])) <= 1e-6
    assert abs(np.angle(p[m // 2, 0]) - 2.0) <= 1e-5


def test_hermite_gaussian_modes_orthogonal_with_power_fractions():
    m, k = 128, 5
    modes = sim.hermite_gaussian_modes(m, k, 30.0).astype(np.complex128)
    g = np.array([[np.vdot(a, b) for b in modes] for a in modes])
    off = np.abs(g - np.diag(np.diag(g)))
    assert off.max() <= 1e-5 * np.abs(np.diag(g)).max()
    pw = np.real(np.diag(g))
    np.testing.assert_allclose(pw / pw.sum(), [0.6, 0.2, 0.1, 0.06, 0.04], rtol=1e-5)


def test_poisson_photon_totals_and_zero_frames():
    rng = np.random.default_rng(2)
    inten = rng.uniform(0, 1, (40, 32, 32))
    inten[3] = 0.0
    for photons in (1e3, 1e5):
        a32 = sim.poisson_amplitudes(inten, photons, rng)
        a = a32.astype(np.float64)
        tot = np.round(a ** 2).sum(axis=(1, 2))
        assert tot[3] == 0.0
        live = np.delete(tot, 3)
        # each total is Poisson(photons): mean within 4 standard errors, spread ~ sqrt(photons)
        assert abs(live.mean() - photons) <= 4 * np.sqrt(photons / live.size)
        assert 0.5 * np.sqrt(photons) <= live.std() <= 1.6 * np.sqrt(photons)
        # amplitudes are the float32 square roots of integer counts
        np.testing.assert_array_equal(np.sqrt(np.round(a ** 2)).astype(np.float32), a32)


def test_simulated_intensities_satisfy_parseval(orc):
    rng = np.random.default_rng(3)
    m = 32
    obj = sim.random_field_object((80, 80), rng)
    probe = sim.gaussian_probe(m, 8.0)
    pos = np.array([[0, 0], [10, 20], [40, 5], [48, 48]], np.float32)
    inten = orc.simulate(obj, probe, pos, squared=True)
    for j, (r, c) in enumerate(pos.astype(int)):
        ex = probe.astype(np.complex128) * obj[r:r + m, c:c + m]
        assert np.isclose(inten[j].sum(), m * m * (np.abs(ex) ** 2).sum(), rtol=1e-10)
