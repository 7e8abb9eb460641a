"""Pins the oracle's solver pieces against closed forms and invariants.

Each check exercises a property fixed by the mathematics rather than the oracle's own
formula: a dropped term, a sign error, a swapped row/column index or a wrong
normalisation makes one of them fail.
"""
import numpy as np
import pytest

from conftest import rel_err


def _rand_c(rng, shape, scale=1.0):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * scale


# ---------------------------------------------------------------- bilinear extraction
def test_bilinear_integer_positions_are_gathers(orc):
    rng = np.random.default_rng(0)
    obj = _rand_c(rng, (40, 52))
    out = orc.extract_bilinear(obj, (7.0, 11.0), 16)
    assert np.array_equal(out, obj[7:23, 11:27])


def test_bilinear_fractional_is_separable_interpolation(orc):
    rng = np.random.default_rng(1)
    obj = _rand_c(rng, (40, 40))
    fx, fy = 0.25, 0.625  # exactly representable: no float rounding in the fractions
    out = orc.extract_bilinear(obj, (5 + fx, 3 + fy), 8)
    rows = (1 - fx) * obj[5:13, :] + fx * obj[6:14, :]
    expect = (1 - fy) * rows[:, 3:11] + fy * rows[:, 4:12]
    assert np.abs(out - expect).max() < 1e-12
    # swapping the roles of the two axes must change the answer (row/col index pin)
    out_sw = orc.extract_bilinear(obj, (5 + fy, 3 + fx), 8)
    assert np.abs(out_sw - expect).max() > 1e-3


def test_bilinear_clamps_at_edges(orc):
    obj = np.arange(16 * 16).reshape(16, 16).astype(complex)
    out = orc.extract_bilinear(obj, (8.5, 8.5), 8)
    assert np.allclose(out[-1, -1], obj[15, 15])  # (15.5, 15.5) clamps to the last pixel


def test_bilinear_weight_quantisation(orc):
    obj = np.zeros((20, 20), complex)
    obj[:, 6] = 1.0
    # fraction 0.3 along columns -> weight quantised to round(0.3*256)/256
    out = orc.extract_bilinear(obj, (0.0, 5.3), 4, quantize=1)
    assert abs(out[0, 0].real - np.round(np.float32(5.3) % 1 * 256) / 256) < 1e-12


# ---------------------------------------------------------------- modulus projection
def test_modulus_projection_closed_form_and_idempotence(orc):
    rng = np.random.default_rng(2)
    m = 32
    psi = _rand_c(rng, (m, m))
    amp = rng.uniform(0.5, 2.0, (m, m))
    out = orc.modulus_projection(psi, amp)
    f = np.fft.fft2(out)
    # |FFT(out)| = A / M^2 * M^2 ... the inverse is unnormalised: FFT(IFFT_u(X)) = M^2 X
    assert np.allclose(np.abs(f), amp, rtol=1e-12)
    assert np.allclose(np.angle(f), np.angle(np.fft.fft2(psi)), atol=1e-10)
    # projecting again with the same amplitudes changes nothing
    assert rel_err(orc.modulus_projection(out, amp), out) < 1e-12


def test_modulus_projection_zero_spectrum_takes_real_amplitude(orc):
    m = 16
    amp = np.ones((m, m)) * 3.0
    out = orc.modulus_projection(np.zeros((m, m), complex), amp)  # atan2(0,0) = 0 -> Phi' = A
    assert np.allclose(np.fft.fft2(out), amp)


# ---------------------------------------------------------------- ePIE
def _epie_problem(rng, m=16, n_side=3, step=5, canvas=None):
    pos = np.array([(i * step, j * step) for i in range(n_side) for j in range(n_side)], np.float32)
    rows = cols = canvas or (int(pos.max()) + m + 2)
    obj = np.exp(1j * rng.uniform(-1, 1, (rows, cols))) * rng.uniform(0.8, 1, (rows, cols))
    probe = np.exp(-((np.arange(m) - m / 2) ** 2)[:, None] / 30 - ((np.arange(m) - m / 2) ** 2)[None, :] / 30)
    probe = probe * np.exp(0.3j * np.arange(m)[None, :])
    return pos, obj, probe[None].astype(complex)


def test_epie_consistent_data_is_a_fixed_point(orc):
    rng = np.random.default_rng(3)
    pos, obj, probe = _epie_problem(rng)
    amp = np.stack([np.abs(np.fft.fft2(obj[int(p[0]):int(p[0]) + 16, int(p[1]):int(p[1]) + 16] * probe[0]))
                    for p in pos])
    st = orc.EpieState(obj, probe, np.max(np.abs(obj) ** 2), 1.0)
    orc.epie_iteration(st, amp, pos, np.arange(len(pos)), update_probe=True)
    assert rel_err(st.obj, obj) < 1e-12
    assert rel_err(st.probe, probe) < 1e-12


def test_epie_zero_amplitudes_closed_form(orc):
    # A = 0  =>  psi = 0  =>  O_r <- O_r - c |P|^2 O_r for a single integer position
    rng = np.random.default_rng(4)
    m = 16
    obj = _rand_c(rng, (24, 24))
    probe = _rand_c(rng, (1, m, m), 0.5)
    pos = np.array([[3.0, 5.0]], np.float32)
    st = orc.EpieState(obj, probe, 1.0, 2.0)
    orc.epie_iteration(st, np.zeros((1, m, m)), pos, [0])
    expect = obj.copy()
    expect[3:19, 5:21] = obj[3:19, 5:21] * (1 - np.abs(probe[0]) ** 2 / 2.0)
    assert rel_err(st.obj, expect) < 1e-13
    assert st.scal[0] == pytest.approx(np.max(np.abs(expect) ** 2), rel=1e-13)


def test_epie_probe_update_uses_updated_object(orc):
    # A = 0: P <- P - (1/objMax) conj(O_new) O_old P on the ROI (ePIE.cpp:117-122)
    rng = np.random.default_rng(5)
    m = 16
    obj = _rand_c(rng, (16, 16))
    probe = _rand_c(rng, (1, m, m), 0.3)
    st = orc.EpieState(obj, probe, 1.0, 4.0)
    orc.epie_iteration(st, np.zeros((1, m, m)), np.zeros((1, 2), np.float32), [0], update_probe=True)
    o_new = obj * (1 - np.abs(probe[0]) ** 2 / 4.0)
    om = np.max(np.abs(o_new) ** 2)
    p_new = probe[0] - np.conj(o_new) * obj * probe[0] / om
    assert rel_err(st.probe[0], p_new) < 1e-12


def test_epie_phase_constraint_clamps_positive_phases(orc):
    rng = np.random.default_rng(6)
    m = 16
    obj = np.exp(1j * rng.uniform(-2, 2, (16, 16)))
    st = orc.EpieState(obj, np.zeros((1, m, m), complex), 1.0, 1.0)
    orc.epie_iteration(st, np.zeros((1, m, m)), np.zeros((1, 2), np.float32), [0], phase_constraint=True)
    assert np.all(np.angle(st.obj) <= 1e-12)
    assert np.allclose(np.abs(st.obj), 1.0)
    neg = np.angle(obj) <= 0
    assert np.allclose(st.obj[neg], obj[neg])


def test_epie_order_matters_but_single_positions_commute_when_disjoint(orc):
    rng = np.random.default_rng(7)
    m = 16
    obj = _rand_c(rng, (16, 40))
    probe = _rand_c(rng, (1, m, m), 0.3)
    amp = rng.uniform(0, 3, (2, m, m))
    pos = np.array([[0, 0], [0, 20]], np.float32)  # non-overlapping ROIs
    a = orc.EpieState(obj, probe, 1.0, 1.0)
    b = orc.EpieState(obj, probe, 1.0, 1.0)
    orc.epie_iteration(a, amp, pos, [0, 1])
    orc.epie_iteration(b, amp, pos, [1, 0])
    assert rel_err(a.obj, b.obj) < 1e-14


# ---------------------------------------------------------------- DM
def test_dm_zero_amplitude_closed_form(orc):
    # A = 0 => Pf = 0 => psi_new = psi - 2 P O_r ; O += s_O conj(P) psi_new (summed over frames)
    rng = np.random.default_rng(8)
    m = 16
    obj = _rand_c(rng, (24, 28))
    probe = _rand_c(rng, (m, m), 0.5)
    pos = np.array([[0, 0], [8, 12]], np.float32)
    psi = _rand_c(rng, (2, m, m), 0.1)
    fac = np.array([0.25, 0.5])
    o, p, s, f, sums = orc.dm_iteration(obj, probe, psi, np.zeros((2, m, m)), pos, fac, update_probe=True)
    eo, ep = obj.copy(), probe.copy()
    es = np.empty_like(psi)
    for j, (r, c) in enumerate(pos.astype(int)):
        po = probe * obj[r:r + m, c:c + m]
        es[j] = psi[j] - 2 * po
    for j, (r, c) in enumerate(pos.astype(int)):
        eo[r:r + m, c:c + m] += 0.5 * np.conj(probe) * es[j]
        ep += 0.25 * np.conj(obj[r:r + m, c:c + m]) * es[j]
    assert rel_err(s, es) < 1e-13
    assert rel_err(o, eo) < 1e-13
    assert rel_err(p, ep) < 1e-13
    assert sums[0] == pytest.approx(np.sum(np.abs(eo) ** 2), rel=1e-12)


def test_dm_consistent_data_from_zero_state(orc):
    # psi = 0 and A = |FFT(P O_r)|: Phi = 2 FFT(P O_r), Pf = P O_r, psi_new = -P O_r,
    # O_r <- O_r - s_O |P|^2 O_r  (one frame, no phase constraint)
    rng = np.random.default_rng(9)
    m = 32
    obj = _rand_c(rng, (40, 40))
    probe = _rand_c(rng, (m, m), 0.5)
    pos = np.array([[3.0, 6.0]], np.float32)
    po = probe * obj[3:3 + m, 6:6 + m]
    amp = np.abs(np.fft.fft2(po))[None]
    o, p, s, f, sums = orc.dm_iteration(obj, probe, np.zeros((1, m, m)), amp, pos, np.array([1.0, 0.125]))
    assert rel_err(s[0], -po) < 1e-12
    eo = obj.copy()
    eo[3:3 + m, 6:6 + m] -= 0.125 * np.abs(probe) ** 2 * obj[3:3 + m, 6:6 + m]
    assert rel_err(o, eo) < 1e-12
    assert rel_err(p, probe) == 0.0  # probe untouched without the update flag


def test_dm_frames_commute(orc):
    # Increments are summed from the iteration-start object: frame order is irrelevant.
    rng = np.random.default_rng(10)
    m = 16
    obj = _rand_c(rng, (30, 30))
    probe = _rand_c(rng, (m, m), 0.5)
    pos = rng.uniform(0, 14, (5, 2)).astype(np.float32)
    psi = _rand_c(rng, (5, m, m), 0.2)
    amp = rng.uniform(0, 5, (5, m, m))
    fac = np.array([0.01, 0.02])
    a = orc.dm_iteration(obj, probe, psi, amp, pos, fac, phase_constraint=True, update_probe=True)
    perm = rng.permutation(5)
    b = orc.dm_iteration(obj, probe, psi[perm], amp[perm], pos[perm], fac, phase_constraint=True, update_probe=True)
    assert rel_err(a[0], b[0]) < 1e-13 and rel_err(a[1], b[1]) < 1e-13
    assert rel_err(a[2][perm], b[2]) < 1e-13


def test_dm_factor_cache_rule(orc):
    # DM.cu:54-73: a cached factor is replaced only when it moved by more than 1e-3.
    # With a zero probe nothing changes, so the passed factors are 1/sum|O|^2 and 1/0.
    rng = np.random.default_rng(11)
    m = 16
    obj = _rand_c(rng, (16, 16), 0.1)
    z = np.zeros((1, m, m))
    zp = np.zeros((1, 2), np.float32)
    so = np.sum(np.abs(obj) ** 2)
    f = orc.dm_iteration(obj, np.zeros((m, m)), z, z, zp, np.array([1 / so + 5e-4, 7.0]))[3]
    assert f[0] == 1 / so + 5e-4          # moved by 5e-4: kept
    assert np.isinf(f[1])                  # 1/0 replaces 7
    f = orc.dm_iteration(obj, np.zeros((m, m)), z, z, zp, np.array([1 / so + 2e-3, 7.0]))[3]
    assert f[0] == pytest.approx(1 / so)  # moved by 2e-3: replaced


# ---------------------------------------------------------------- probe init / simulator
def test_probe_init_properties(orc):
    rng = np.random.default_rng(12)
    m = 64
    amp = rng.uniform(0, 10, (4, m, m))
    probe, pmax = orc.probe_init(amp, 2, beam_size=100e-9, dx_s=10e-9)
    assert np.max(np.abs(probe[0])) == pytest.approx(1.0)
    assert pmax == pytest.approx(1.0)
    assert np.all(probe[1] == 0)
    # spectrum magnitude is proportional to the mean amplitude (modulus projection)
    f = np.abs(np.fft.fft2(probe[0]))
    avg = amp.mean(axis=0)
    ratio = f / avg
    assert np.allclose(ratio, ratio.mean(), rtol=1e-9)


def test_simulate_matches_numpy_at_integer_positions(orc):
    rng = np.random.default_rng(13)
    m = 32
    obj = _rand_c(rng, (50, 60))
    probe = _rand_c(rng, (m, m))
    pos = np.array([[0, 0], [10, 20], [18, 28]], np.float32)
    out = orc.simulate(obj, probe, pos, squared=True)
    for j, (r, c) in enumerate(pos.astype(int)):
        ref = np.abs(np.fft.fft2(obj[r:r + m, c:c + m] * probe)) ** 2
        assert rel_err(out[j], ref) < 1e-12
