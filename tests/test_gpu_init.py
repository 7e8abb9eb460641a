"""Probe initialisation paths on the GPU against the FP64 oracle.

Probe::orthogonalize (Probe.cpp:228-242) on random complex modes, Probe::initProbeModes
(Probe.cpp:320-330) through ptyk_begin_modal, and Probe::initMLH (Probe.cpp:376-494) through the
driver's ptycho_mls_probe_modes with fixed draws.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _unit(shape, rng):
    u = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return (u / np.sqrt(np.mean(np.abs(u) ** 2))).astype(np.complex64)


@pytest.mark.parametrize("k,m", [(2, 32), (4, 64), (5, 256)])
def test_orthogonalize_matches_oracle(gpu, orc, k, m):
    from ptychopy_amd._engine import Engine
    rng = np.random.default_rng(k * 100 + m)
    # well-conditioned modes: a common base plus independent parts
    base = _unit((m, m), rng)
    modes = np.stack([base + 0.7 * _unit((m, m), rng) for _ in range(k)]).astype(np.complex64)
    print(f"orthogonalize k={k} m={m}", flush=True)
    with Engine(4, m, (m + 8, m + 8), n_modes=k) as e:
        e.set_probe(modes)
        e.orthogonalize_modes()
        got = e.get_probe()
    ref = orc.orthogonalize_modes(modes.astype(np.complex128))
    for i in range(k):
        assert rel_err(got[i], ref[i]) < 1e-5, f"mode {i}"


def test_begin_modal_matches_oracle_init_modes(gpu, orc):
    """Mode 0 is untouched; the modal probe maximum is max_q (|P0|^2 + (K-1) |0.05 P0|^2) from the
    stored intensities; mode 1 = 0.05 P0 minus its projection on P0, i.e. zero up to FP32 rounding
    (modes >= 2 are fixed by that rounding residue in the reference too: not compared)."""
    from ptychopy_amd._engine import Engine, SCAL_PROBE_MAX
    rng = np.random.default_rng(3)
    k, m = 3, 64
    p0 = _unit((m, m), rng)
    probe = np.zeros((k, m, m), np.complex64)
    probe[0] = p0
    with Engine(4, m, (m + 8, m + 8), n_modes=k) as e:
        e.set_probe(probe)
        e.refresh_maxima()
        e.begin_modal()
        got = e.get_probe()
        pmax = e.get_scalars()[SCAL_PROBE_MAX]
    ref, inten = orc.init_probe_modes(probe.astype(np.complex128))
    assert np.array_equal(got[0], p0)
    inten[0] = np.abs(p0.astype(np.complex128)) ** 2
    assert pmax == pytest.approx(inten.sum(axis=0).max(), rel=1e-6)
    assert np.linalg.norm(got[1]) < 1e-5 * np.linalg.norm(p0)
    assert np.linalg.norm(ref[1]) < 1e-12 * np.linalg.norm(p0)


@pytest.mark.parametrize("k,m", [(1, 64), (4, 64), (5, 256)])
def test_mls_initial_modes_match_oracle(gpu, orc, k, m):
    from ptychopy_amd import _driver
    from ptychopy_amd._engine import Engine
    lambd, dx_s, beam = 2.48e-10, 5.631549e-09, 1e-7
    draws = np.random.default_rng(k + m).standard_normal(2 * (k - 1)).astype(np.float32)
    print(f"initMLH k={k} m={m}", flush=True)
    with Engine(1, m, (m, m), n_modes=k) as e:
        got = _driver.mls_probe_modes(e, k, lambd, dx_s, beam, draws)
    ref = orc.mls_probe_modes(m, k, lambd, dx_s, beam, draws.astype(np.float64))
    assert rel_err(got, ref) < 1e-5
    for i in range(k):
        assert rel_err(got[i], ref[i]) < 1e-5, f"mode {i}"
