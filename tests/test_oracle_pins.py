"""Closed-form pins of the oracle branches the solver-level tests leave open.

Each expected value below is derived by hand from the reference's equations for inputs chosen so
the answer has a closed form (constant spectra, delta probes, orthogonal patterns, hand-picked
coefficients); a dropped term, a missing conjugate, a wrong normalisation or a swapped axis moves
the result by O(1).  References:
  modal modulus             utilitiesKernels.cu:655-682 (d_adjustModalFFT)
  ePIE Fourier error        utilitiesKernels.cu:684-712 (d_calculateER), ePIE.cpp:124-128
  MLs Fourier error         Diffractions.cpp:261-279, utilitiesKernels.cu:1672-1678
  variable probe            Probe.cpp:1366-1477
  mode directions           MLS.cpp:251-297
  calc_object_norm          Probe.cpp:671-699
  remove_extra_degree       Probe.cpp:615-669
  orthogonalize / modes     Probe.cpp:228-242, 320-330; ProbeKernels.cu:82-97, 509-515
  DM increment clamp        SampleKernels.cu:468-476
"""
import numpy as np
import pytest

from conftest import rel_err


def _unit(shape, rng):
    u = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return u / np.sqrt(np.mean(np.abs(u) ** 2))


def _orth_pair(m, rng):
    """Two m x m patterns with unit rms and mean(U conj W) = 0."""
    u = _unit((m, m), rng)
    w = _unit((m, m), rng)
    w = w - np.mean(w * np.conj(u)) / np.mean(np.abs(u) ** 2) * u
    return u, w / np.sqrt(np.mean(np.abs(w) ** 2))


# ------------------------------------------------------------------------------- ePIE / modulus
def test_modal_modulus_magnitude_phase_idempotence(orc):
    rng = np.random.default_rng(1)
    k, m = 3, 16
    psi = rng.standard_normal((k, m, m)) + 1j * rng.standard_normal((k, m, m))
    amp = rng.uniform(0.5, 2.0, (m, m))
    out = orc.modulus_modes(psi, amp)
    phi = np.fft.fft2(psi)
    f = np.fft.fft2(out)
    # sum_k |Phi'_k|^2 = A^2 (the unnormalised inverse and the forward transform cancel the 1/M^2)
    assert np.allclose(np.sum(np.abs(f) ** 2, axis=0), amp ** 2, rtol=1e-10)
    # every mode keeps its phase, and all modes of a pixel share one real positive factor
    ratio = f / phi
    assert np.allclose(ratio.imag, 0, atol=1e-10) and np.all(ratio.real > 0)
    assert np.allclose(ratio, ratio[0:1], rtol=1e-10)
    # a projection: applying it again changes nothing
    assert rel_err(orc.modulus_modes(out, amp), out) < 1e-12
    # one mode: the single-mode projection A exp(i arg Phi)
    assert rel_err(orc.modulus_modes(psi[:1], amp)[0], orc.modulus_projection(psi[0], amp)) < 1e-12


@pytest.mark.parametrize("k,c,a", [(2, (1.0, 2.0), 4.0), (1, (2.0,), 5.0), (2, (1.0, 2.0j), 1.0)])
def test_epie_fourier_error_closed_form(orc, k, c, a):
    """O = 1 and P_k = c_k delta: Phi_k = c_k at every frequency, so the error of the one
    position is sum |A - |sum_k c_k|| / K / sum A = |a - |sum c|| / (K a): a coherent sum over
    modes, divided by the mode count and by the sum of the amplitudes."""
    m = 8
    obj = np.ones((m + 4, m + 4), np.complex128)
    probe = np.zeros((k, m, m), np.complex128)
    for i, ci in enumerate(c):
        probe[i, 0, 0] = ci
    amp = np.full((1, m, m), a)
    st = orc.EpieState(obj, probe, 1.0, 1.0)
    err = orc.epie_iteration(st, amp, np.array([[2.0, 2.0]], np.float32), [0], k_use=k, modal=k > 1, rms=True)
    assert err == pytest.approx(abs(a - abs(sum(c))) / (k * a), rel=1e-12)


def test_mls_fourier_error_and_modulus_closed_form(orc):
    """Constant spectra c_k and amplitude a: ap = sqrt(sum |c_k|^2), error = sqrt(mean((a - ap)^2
    / 0.25)) = 2 |a - ap|, chi_k = (a / (ap + 1e-9) - 1) c_k delta."""
    m = 8
    c = np.array([3.0, 4.0])
    spec = np.broadcast_to(c[:, None, None], (2, m, m)).astype(np.complex128)
    chi, err = orc.mls_frame_modulus(spec, np.full((m, m), 2.0))
    assert err == pytest.approx(2 * abs(2.0 - 5.0), rel=1e-12)  # 6: not 3 (plain RMS), not 10 (coherent)
    r = 2.0 / (5.0 + 1e-9) - 1.0
    expect = np.zeros((2, m, m), np.complex128)
    expect[:, 0, 0] = r * c
    assert np.allclose(chi, expect, atol=1e-12)


# ------------------------------------------------------------------------------- MLs pieces
@pytest.mark.parametrize("big", [False, True])
def test_mls_variable_probe_closed_form(orc, big):
    """Two frames (of n = 4) with probe updates pu_j = a_j U + (b_j + g) W, mean update g W,
    V = U, and weights p / max p = 1/2 (the canvas maximum lies outside both ROIs).
    img_j = (a_j U + b_j W) / 2; c_j = (Re(a_j) / 2 + e_j) / (e_1^2 + e_2^2); the update
    (alpha U + beta W) / M^2 with alpha = sum c_j a_j / 2, beta = sum c_j b_j / 2 is scaled by
    (G / n) / max(1, its rms) and added to V, which is renormalised to unit rms.  Then with
    op_j = o_j and chi_j = d_j o_j V_new: e_j += d_j |o_j|^2 / (|o_j|^2 + 0.1 mean |o|^2)."""
    rng = np.random.default_rng(5)
    m, n = 8, 4
    mm = m * m
    U, Wp = _orth_pair(m, rng)
    s = 400.0 if big else 1.0  # big: the update's rms exceeds 1 and the max(1, .) clamp binds
    a = np.array([1.5 + 0.5j, -0.7 + 2.0j]) * s
    b = np.array([0.3 - 1.1j, 2.2 + 0.4j]) * s
    g = 0.6 - 0.2j
    e = np.array([0.8, -0.5, 9.0, 9.0])
    o = np.array([1.3 - 0.4j, 0.6 + 0.9j])
    d = np.array([0.25, -1.5])
    rows, cols = 20, 20
    roi = np.array([[1, 2], [5, 6], [0, 0], [0, 0]], np.int32)
    illum = np.full((rows, cols), 2.0)
    illum[19, 19] = 4.0  # the maximum, outside both ROIs: w = 1/2 inside them
    frames = [0, 1]
    pu = np.stack([a[j] * U + (b[j] + g) * Wp for j in range(2)])
    pum = g * Wp
    # expected V from the closed form
    evn = e[0] ** 2 + e[1] ** 2
    cj = (a.real / 2 + e[:2]) / evn
    alpha = np.sum(cj * a / 2) / mm
    beta = np.sum(cj * b / 2) / mm
    rms_acc = np.sqrt(abs(alpha) ** 2 + abs(beta) ** 2)
    sc = (2 / n) / max(1.0, rms_acc)
    v_new = (1 + alpha * sc) * U + beta * sc * Wp
    v_new /= np.sqrt(abs(1 + alpha * sc) ** 2 + abs(beta * sc) ** 2)
    if big:
        assert rms_acc > 1
    op = np.stack([np.full((m, m), o[j]) for j in range(2)])
    chi0 = np.stack([d[j] * o[j] * v_new for j in range(2)])
    v, e1 = orc.mls_variable_probe(frames, roi, illum, pu, pum, op, chi0, U, e)
    assert rel_err(v, v_new) < 1e-10
    den = np.abs(o) ** 2
    assert np.allclose(e1[:2], e[:2] + d * den / (den + 0.1 * den.mean()), rtol=1e-9)
    assert np.array_equal(e1[2:], e[2:])


def test_mls_mode_directions_closed_form(orc):
    """Mode 2 of two, one frame, op = o: pum = w c conj(o) U with w = p / (p + 0.01 MAX_ILLUM);
    its object direction is P_2 chi = p c W U, without the conjugate of P_2 (MLS.cpp:294-297).
    Mode 1 without the sub-pixel shift: chi conj(vp)."""
    rng = np.random.default_rng(6)
    m = 8
    U, W = _unit((m, m), rng), _unit((m, m), rng)
    Y = _unit((m, m), rng)
    c, o, p, v = 0.7 - 1.2j, 1.1 + 0.6j, -0.4 + 0.9j, 0.5 + 0.5j
    illum = np.full((12, 12), 3.0)
    max_illum = 50.0
    w = 3.0 / (3.0 + 0.01 * max_illum)
    probe = np.stack([_unit((m, m), rng), p * W])
    roi = np.array([[2, 3]], np.int32)
    sub = np.zeros((1, 2), np.float32)
    chi = (c * U)[None]
    op = np.full((1, m, m), o)
    vp = (v * Y)[None]
    pu, pum, oup = orc.mls_mode_directions(2, [0], roi, sub, illum, max_illum, chi, op, vp, probe)
    assert rel_err(pu[0], c * np.conj(o) * U) < 1e-12
    assert rel_err(pum, w * c * np.conj(o) * U) < 1e-12
    assert rel_err(oup[0], p * W * c * U) < 1e-12
    assert rel_err(oup[0], np.conj(p * W) * c * U) > 0.1  # the conjugated form is a different direction
    _, _, oup1 = orc.mls_mode_directions(1, [0], roi, sub, illum, max_illum, chi, op, vp, probe)
    assert rel_err(oup1[0], c * U * np.conj(v * Y)) < 1e-12


def test_mls_object_norm_closed_form(orc):
    """Box rows split in two halves with |O| = r1, r2 and illumination w1, w2:
    object_norm = sqrt((r1^2 w1^2 + r2^2 w2^2) / (w1^2 + w2^2)) (weights squared, not linear);
    probe and variable mode are multiplied by it, the whole object divided by it."""
    rng = np.random.default_rng(7)
    rows, cols, m = 8, 10, 4
    r1, r2, w1, w2 = 2.0, 0.5, 3.0, 1.0
    obj = _unit((rows, cols), rng)
    illum = rng.uniform(0.1, 9.0, (rows, cols))
    box = (2, 5, 3, 6)  # 1-based inclusive: rows 1..4, cols 2..5
    obj[1:3, 2:6] = r1 * np.exp(1j * rng.uniform(-3, 3, (2, 4)))
    obj[3:5, 2:6] = r2 * np.exp(1j * rng.uniform(-3, 3, (2, 4)))
    illum[1:3, 2:6] = w1
    illum[3:5, 2:6] = w2
    probe = _unit((2, m, m), rng)
    vmode = _unit((m, m), rng)
    onorm = np.sqrt((r1 ** 2 * w1 ** 2 + r2 ** 2 * w2 ** 2) / (w1 ** 2 + w2 ** 2))
    o2, p2, v2 = orc.mls_object_norm(box, obj, illum, probe, vmode)
    assert rel_err(o2, obj / onorm) < 1e-12
    assert rel_err(p2, probe * onorm) < 1e-12
    assert rel_err(v2, vmode * onorm) < 1e-12


def test_mls_remove_extra_degree_closed_form(orc):
    """n = 3, e0 = (2, 4, 7), e1 = (1, 3, 5), V = 2 U (rms 2): V -> U, e1 -> 2 e1 = (2, 6, 10);
    over the first two positions only: means 3 and 4, e0 -> ((2-3) 0.99 + 1, (4-3) 0.99 + 1) =
    (0.01, 1.99), e1 -> (-2, 2); cn = rms(0.01, 1.99); e0[:2] /= cn, P0 *= cn; position 3 keeps
    e0 = 7 and e1 = 10."""
    rng = np.random.default_rng(8)
    m = 4
    U = _unit((m, m), rng)
    probe = _unit((2, m, m), rng)
    evo = np.array([[2.0, 4.0, 7.0], [1.0, 3.0, 5.0]])
    p2, v2, e2 = orc.mls_remove_extra(probe, 2 * U, evo)
    cn = np.sqrt((0.01 ** 2 + 1.99 ** 2) / 2)
    assert rel_err(v2, U) < 1e-12
    assert np.allclose(e2[0], [0.01 / cn, 1.99 / cn, 7.0], rtol=1e-12)
    assert np.allclose(e2[1], [-2.0, 2.0, 10.0], rtol=1e-12)
    assert rel_err(p2[0], probe[0] * cn) < 1e-12
    assert np.array_equal(p2[1], probe[1])


# ------------------------------------------------------------------------------- probe modes
def test_orthogonalize_real_modes_is_gram_schmidt(orc):
    """For real modes the reference's coefficient equals the textbook one, so mode i becomes v_i
    minus its projection on the span of v_0..v_{i-1} (from a QR factorisation)."""
    rng = np.random.default_rng(9)
    k, m = 4, 6
    v = rng.standard_normal((k, m, m)).astype(np.complex128)
    out = orc.orthogonalize_modes(v)
    a = v.reshape(k, -1).real.T
    q, _ = np.linalg.qr(a)
    for i in range(k):
        qi = q[:, :i]
        expect = a[:, i] - qi @ (qi.T @ a[:, i])
        assert rel_err(out[i].ravel(), expect) < 1e-10
    g = out.reshape(k, -1) @ out.reshape(k, -1).conj().T
    assert np.allclose(g - np.diag(np.diag(g)), 0, atol=1e-9 * np.abs(g).max())


def test_orthogonalize_complex_coefficient_is_conjugated(orc):
    """v_1 = c v_0 with complex c: the coefficient sum conj(v_1) v_0 / sum |v_0|^2 is conj(c),
    so mode 1 becomes (c - conj(c)) v_0 = 2 i Im(c) v_0 (the textbook projection would give 0)."""
    rng = np.random.default_rng(10)
    m = 6
    u0 = _unit((m, m), rng)
    c = 0.3 + 0.8j
    out = orc.orthogonalize_modes(np.stack([u0, c * u0]))
    assert rel_err(out[0], u0) == 0
    assert rel_err(out[1], 2j * c.imag * u0) < 1e-12


def test_init_probe_modes_from_base(orc):
    """initProbeModes: modes >= 1 start as 0.05 P0 with intensities |0.05 P0|^2; the real factor
    makes the conjugated coefficient exact, so orthogonalisation leaves mode 1 at zero."""
    rng = np.random.default_rng(11)
    m = 8
    p0 = _unit((m, m), rng)
    probe = np.stack([p0, _unit((m, m), rng), _unit((m, m), rng)])
    out, inten = orc.init_probe_modes(probe)
    assert np.array_equal(out[0], p0)
    assert np.allclose(inten[1], np.abs(0.05 * p0) ** 2, rtol=1e-12)
    assert np.allclose(inten[2], np.abs(0.05 * p0) ** 2, rtol=1e-12)
    assert np.linalg.norm(out[1]) < 1e-12 * np.linalg.norm(p0)


# ------------------------------------------------------------------------------- DM clamp
@pytest.mark.parametrize("theta", [0.7, -0.7])
def test_dm_phase_clamp_acts_on_the_increment(orc, theta):
    """One frame, psi = 0, A = 0: psi_new = -2 P O, the object increment u = -2 s |P|^2 O has phase
    theta + pi.  theta = 0.7: arg u = 0.7 - pi < 0, unclamped, O -> O (1 - 2 s |P|^2) (clamping O
    itself would make it real).  theta = -0.7: arg u = pi - 0.7 > 0, clamped to |u|, O -> O +
    2 s |P|^2 (no clamp would shrink O)."""
    rng = np.random.default_rng(12)
    m = 8
    rows = cols = m + 4
    obj = np.full((rows, cols), np.exp(1j * theta))
    probe = 0.3 * _unit((m, m), rng)
    s_o = 0.7
    o, _, _, _, _ = orc.dm_iteration(obj, probe, np.zeros((1, m, m)), np.zeros((1, m, m)),
                                     np.array([[2.0, 1.0]], np.float32), np.array([0.0, s_o]), phase_constraint=True)
    p2 = np.abs(probe) ** 2
    expect = obj.copy()
    roi = expect[2:2 + m, 1:1 + m]
    if theta > 0:
        expect[2:2 + m, 1:1 + m] = roi * (1 - 2 * s_o * p2)
    else:
        expect[2:2 + m, 1:1 + m] = roi + 2 * s_o * p2
    assert rel_err(o, expect) < 1e-12


# ------------------------------------------------------------------------------- initMLH
def _annulus_pixels(m, lambd, dx_s):
    """Grid pixels with D_H/2 <= r <= D_FZP/2 (30 um .. 90 um) on the zone-plate grid."""
    fl = 2 * 90e-6 * 50e-9 / lambd
    dx = lambd * fl / (m * dx_s)
    x = -dx * m / 2 + dx * m / (m - 1) * np.arange(m)
    r = np.sqrt(x[None, :] ** 2 + x[:, None] ** 2)
    return int(np.count_nonzero((r <= 90e-6) & (r >= 30e-6)))


@pytest.mark.parametrize("m", [64, 128])
def test_mls_zone_plate_parseval(orc, m):
    """The zone-plate kernel has unit modulus on the annulus and zero elsewhere, the far-field
    factor has unit modulus, and the forward FFT is unnormalised: sum |P0|^2 = M^2 x (annulus pixels)."""
    lambd, dx_s = 2.48e-10, 5.631549e-09
    p0 = orc.mls_zone_plate_probe(m, lambd, dx_s, 1e-7)
    n_ann = _annulus_pixels(m, lambd, dx_s)
    assert n_ann > 10
    assert np.sum(np.abs(p0) ** 2) == pytest.approx(m * m * n_ann, rel=1e-12)


def test_mls_initial_modes_orthogonal_strongest_first(orc):
    """The eigenmodes are mutually orthogonal with powers in decreasing order, and the unitary
    mixing keeps the incoherent sum: sum_a |P_a|^2 = K |P0|^2 at every pixel (each tilted copy has
    the modulus of mode 0)."""
    m, k = 64, 4
    draws = [0.3, -0.8, 1.1, 0.4, -1.7, 0.2]
    p0 = orc.mls_zone_plate_probe(m, 2.48e-10, 5.631549e-09, 1e-7)
    modes = orc.mls_probe_modes(m, k, 2.48e-10, 5.631549e-09, 1e-7, draws)
    g = modes.reshape(k, -1).conj() @ modes.reshape(k, -1).T
    off = g - np.diag(np.diag(g))
    assert np.abs(off).max() < 1e-9 * np.abs(g).max()
    pw = np.real(np.diag(g))
    assert np.all(np.diff(pw) < 0)
    assert np.allclose(np.sum(np.abs(modes) ** 2, axis=0), k * np.abs(p0) ** 2, rtol=1e-9, atol=1e-9 * pw[0])
    # each eigenvector's largest component is real positive: the mode's overlap with the tilted
    # copy it draws most from is real positive
    tilts = np.stack([p0] + [p0 * np.exp(-2j * np.pi * (draws[2 * i - 2] * (np.arange(m) / m - 0.5)[None, :] +
                                                        draws[2 * i - 1] * (np.arange(m) / m - 0.5)[:, None]))
                             for i in range(1, k)])
    for a in range(k):
        ov = tilts.reshape(k, -1).conj() @ modes[a].ravel()
        big = int(np.argmax(np.abs(ov)))
        assert abs(ov[big].imag) < 1e-9 * abs(ov[big]) and ov[big].real > 0
