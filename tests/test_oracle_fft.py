"""Pins the oracle's FP64 FFT (oracle.cpp orc_fft2) against numpy.fft and closed forms."""
import numpy as np
import pytest

from conftest import rel_err


@pytest.mark.parametrize("m", [16, 32, 64, 128, 256, 512])
def test_fft2_matches_numpy(orc, m):
    rng = np.random.default_rng(m)
    x = rng.standard_normal((3, m, m)) + 1j * rng.standard_normal((3, m, m))
    assert rel_err(orc.fft2(x), np.fft.fft2(x)) < 1e-13
    # inverse is unnormalised (cuFFT CUFFT_INVERSE convention)
    assert rel_err(orc.fft2(x, inverse=True), np.fft.ifft2(x) * m * m) < 1e-13


def test_fft2_closed_forms(orc):
    m = 64
    d = np.zeros((m, m), complex)
    d[0, 0] = 1
    assert np.allclose(orc.fft2(d), np.ones((m, m)))          # delta -> constant
    k = (5, 9)
    i, j = np.mgrid[0:m, 0:m]
    tone = np.exp(2j * np.pi * (k[0] * i + k[1] * j) / m)
    f = orc.fft2(tone)
    expect = np.zeros((m, m), complex)
    expect[k] = m * m                                        # single tone -> delta
    assert np.abs(f - expect).max() < 1e-9
    rng = np.random.default_rng(1)
    x = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    # Parseval and round trip
    assert abs(np.sum(np.abs(orc.fft2(x)) ** 2) / (m * m) - np.sum(np.abs(x) ** 2)) < 1e-9 * np.sum(np.abs(x) ** 2)
    assert rel_err(orc.fft2(orc.fft2(x), inverse=True) / (m * m), x) < 1e-14
