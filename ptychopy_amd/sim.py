"""Seeded synthetic ptychography data (objects, probes, scans, Poisson frames).

The object/probe/scan statistics follow the benchmark workloads:
  * object: amplitude U[0.8, 1.0] smoothed by a Gaussian (sigma 4 px), phase a Gaussian
    random field (white noise filtered to a 6 px correlation length, rescaled to a given std);
  * probe: Gaussian amplitude of a given FWHM with a quadratic (defocus) phase reaching a
    given value at the array edge;
  * scan: raster with fractional pixel steps plus optional uniform jitter.
Diffraction amplitudes are formed on the GPU by the engine (``Engine.simulate``:
|FFT2(bilinear(O, pos) * P)|^2), scaled to a photon budget and Poisson-sampled here.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage

SEED_BASE = 20261202


def rng_for(config_id: int, purpose: int = 0) -> np.random.Generator:
    """One independent stream per (config, purpose)."""
    return np.random.default_rng([SEED_BASE + config_id, purpose])


def random_field_object(shape, rng, amp_lo=0.8, amp_hi=1.0, amp_sigma=4.0, phase_corr=6.0,
                        phase_std=0.5) -> np.ndarray:
    amp = rng.uniform(amp_lo, amp_hi, size=shape)
    amp = ndimage.gaussian_filter(amp, amp_sigma, mode="wrap")
    ph = rng.standard_normal(size=shape)
    ph = ndimage.gaussian_filter(ph, phase_corr, mode="wrap")
    ph = (ph - ph.mean()) / max(ph.std(), 1e-30) * phase_std
    return (amp * np.exp(1j * ph)).astype(np.complex64)


def gaussian_probe(m: int, fwhm: float, defocus_edge: float = 2.0) -> np.ndarray:
    """Gaussian amplitude (FWHM in px) with phase defocus_edge * (r / (m/2))^2."""
    y, x = np.mgrid[0:m, 0:m].astype(np.float64)
    r2 = (x - m / 2) ** 2 + (y - m / 2) ** 2
    sigma = fwhm / (2.0 * np.sqrt(2.0 * np.log(2.0)))
    amp = np.exp(-r2 / (2 * sigma * sigma))
    ph = defocus_edge * r2 / (m / 2) ** 2
    return (amp * np.exp(1j * ph)).astype(np.complex64)


def hermite_gaussian_modes(m: int, k: int, fwhm: float, powers=(0.6, 0.2, 0.1, 0.06, 0.04),
                           defocus_edge: float = 2.0) -> np.ndarray:
    """k incoherent probe modes: Hermite-Gaussian HG_ab (a + b = 0, 1, 1, 2, 2, ...) of the given
    fundamental FWHM with a common defocus phase, orthogonalised (Gram-Schmidt on the grid) and
    scaled to the power fractions (total power that of the unit-peak fundamental)."""
    from numpy.polynomial.hermite import hermval
    y, x = np.mgrid[0:m, 0:m].astype(np.float64)
    x, y = x - m / 2, y - m / 2
    sigma = fwhm / (2.0 * np.sqrt(2.0 * np.log(2.0)))
    w = np.sqrt(2.0) * sigma
    g = np.exp(-(x * x + y * y) / (w * w))
    order = [(a, n - a) for n in range(8) for a in range(n + 1)][:k]
    ph = np.exp(1j * defocus_edge * (x * x + y * y) / (m / 2) ** 2)
    modes = []
    for a, b in order:
        v = hermval(np.sqrt(2) * x / w, [0] * a + [1]) * hermval(np.sqrt(2) * y / w, [0] * b + [1]) * g * ph
        for u in modes:
            v = v - np.vdot(u, v) / np.vdot(u, u) * u
        modes.append(v)
    tot = float(np.sum(g * g))
    pw = np.asarray(powers[:k], np.float64)
    pw = pw / pw.sum()
    out = np.array([mo * np.sqrt(pw[i] * tot / np.sum(np.abs(mo) ** 2)) for i, mo in enumerate(modes)])
    return out.astype(np.complex64)


def raster_positions(n_side: int, step: float, rng=None, jitter: float = 0.0, origin=(0.0, 0.0)) -> np.ndarray:
    """(n_side^2, 2) float32 (row, col) positions, row-major scan order, min at ``origin``."""
    i, j = np.meshgrid(np.arange(n_side), np.arange(n_side), indexing="ij")
    pos = np.stack([i.ravel() * step, j.ravel() * step], axis=1)
    if jitter > 0 and rng is not None:
        pos = pos + rng.uniform(-jitter, jitter, size=pos.shape)
    pos = pos - pos.min(axis=0) + np.asarray(origin, np.float64)
    return pos.astype(np.float32)


def canvas_for(pos: np.ndarray, m: int) -> tuple[int, int]:
    """Canvas = ceil(max - min + M) per axis (IPtychoScanMesh::calculateMeshDimensions)."""
    ext = pos.max(axis=0).astype(np.float64) - np.minimum(pos.min(axis=0), 0).astype(np.float64) + m
    return int(np.ceil(ext[0])), int(np.ceil(ext[1]))


def poisson_amplitudes(intensity: np.ndarray, photons: float, rng) -> np.ndarray:
    """Scale each frame to ``photons`` total counts, draw Poisson counts, return sqrt(counts)."""
    it = np.asarray(intensity, np.float64)
    tot = it.reshape(it.shape[0], -1).sum(axis=1)
    scale = np.where(tot > 0, photons / np.maximum(tot, 1e-300), 0.0)
    counts = rng.poisson(it * scale[:, None, None])
    return np.sqrt(counts).astype(np.float32)
