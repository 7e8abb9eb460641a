"""Pins of the FP64 MLs oracle (oracle/oracle_mls.cpp) by closed forms, invariants and
independent numpy computations -- each chosen so that a dropped term, a wrong sign/conjugate or
a transposed axis fails one of them."""
import numpy as np
import pytest

from problems import make_mls_problem, mls_exit_amplitudes, mls_groups


def _field(rng, m):
    return rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))


def test_shift_integer_is_roll(orc):
    rng = np.random.default_rng(1)
    x = _field(rng, 16)
    # +1 column, -2 rows: the probe moves towards +sub (Probe::shift_probe phase e^{-2 pi i s f})
    y = orc.mls_shift(x, 1.0, -2.0)
    assert np.allclose(y, np.roll(np.roll(x, 1, axis=1), -2, axis=0), atol=1e-12)


def test_shift_is_unitary_and_invertible(orc):
    rng = np.random.default_rng(2)
    x = _field(rng, 32)
    y = orc.mls_shift(x, 0.37, -0.21)
    assert np.isclose(np.linalg.norm(y), np.linalg.norm(x), rtol=1e-12)
    z = orc.mls_shift(y, -0.37, 0.21)
    assert np.allclose(z, x, atol=1e-11)


def test_position_derivatives_plane_wave(orc):
    m = 32
    r, c = np.mgrid[0:m, 0:m]
    a, b = 3, -5
    x = np.exp(2j * np.pi * (a * c + b * r) / m)
    _, _, dX, dY = orc.mls_position_step(x)
    # dX is d/d(column), dY is d/d(row) (d_multiplyRow / d_multiplyColum)
    assert np.allclose(dX, 2j * np.pi * a / m * x, atol=1e-10)
    assert np.allclose(dY, 2j * np.pi * b / m * x, atol=1e-10)


def test_position_step_recovers_linear_residual(orc):
    rng = np.random.default_rng(3)
    m = 32
    op = np.exp(1j * rng.uniform(-1, 1, (m, m)))
    vp = _field(rng, m)
    _, _, dX, dY = orc.mls_position_step(op)
    chi = 0.13 * dX * vp
    dx, dy, _, _ = orc.mls_position_step(op, vp, chi)
    assert np.isclose(dx, 0.13, rtol=1e-10)
    ay, ax = dY * vp, dX * vp
    assert np.isclose(dy, 0.13 * np.sum((ax * np.conj(ay)).real) / np.sum(np.abs(ay) ** 2), rtol=1e-9)


def test_lsq_closed_form_orthogonal_directions(orc):
    """Identical frames, dOP orthogonal to dPO, chi = a dOP + b dPO: the 0.1 mean regulariser
    scales the diagonal by 1.1, so the steps are (a, b) / 1.1 (lambda = FLT_EPSILON negligible)."""
    rng = np.random.default_rng(4)
    m, g = 16, 3
    op = np.ones((g, m, m), np.complex128)
    vp = np.ones((g, m, m), np.complex128)
    half = np.zeros((m, m))
    half[: m // 2] = 1.0
    oup = np.broadcast_to(half * (2.0 + 1.0j), (g, m, m)).copy()
    pum = (1.0 - half) * (0.5 - 3.0j)
    a, b = 0.7 - 0.2j, -1.3 + 0.4j
    chi = a * oup * vp + b * pum * op
    x = orc.mls_lsq(oup, vp, pum, op, chi)
    # real right-hand side: Re(conj(d) a d) = Re(a)|d|^2
    assert np.allclose(x[:, 0], a.real / 1.1, rtol=1e-6)
    assert np.allclose(x[:, 1], b.real / 1.1, rtol=1e-6)


def test_lsq_matches_numpy_solve(orc):
    rng = np.random.default_rng(5)
    m, g = 16, 4
    oup, vp, op, chi = (np.stack([_field(rng, m) for _ in range(g)]) for _ in range(4))
    pum = _field(rng, m)
    x = orc.mls_lsq(oup, vp, pum, op, chi)
    dop, dpo = oup * vp, pum * op
    lam = np.float32(np.finfo(np.float32).eps / (m * m)) * m * m
    a1 = np.sum(np.abs(dop) ** 2, axis=(1, 2)) + lam
    a4 = np.sum(np.abs(dpo) ** 2, axis=(1, 2)) + lam
    a2 = np.sum(dop * np.conj(dpo), axis=(1, 2))
    b1 = np.sum((np.conj(dop) * chi).real, axis=(1, 2))
    b2 = np.sum((np.conj(dpo) * chi).real, axis=(1, 2))
    for j in range(g):
        A = np.array([[a1[j] + 0.1 * a1.mean(), a2[j]], [np.conj(a2[j]), a4[j] + 0.1 * a4.mean()]])
        assert np.allclose(x[j], np.linalg.solve(A, [b1[j], b2[j]]), rtol=1e-9)


def test_ortho_modes_is_gram_schmidt_then_power_order(orc):
    rng = np.random.default_rng(6)
    k, m = 4, 16
    P = np.stack([_field(rng, m) * s for s in (3.0, 0.5, 2.0, 1.0)])
    out = orc.mls_ortho_modes(P)
    # Gram-Schmidt without normalisation = Q_i R_ii of the QR factorisation
    Q, R = np.linalg.qr(P.reshape(k, -1).T)
    gs = (Q * np.diag(R)).T.reshape(k, m, m)
    pw = [np.sum(np.abs(P[i]) ** 2) for i in range(k)]
    order = sorted(range(k), key=lambda i: (-pw[i], i))
    assert np.allclose(out[0], gs[0])
    for i in range(1, k):
        assert np.allclose(out[i], gs[order[i]], atol=1e-10)


def test_roi_origins_are_rounded_positions(orc):
    rng = np.random.default_rng(7)
    m, rows, cols = 16, 101, 96
    pos = rng.uniform(-20, 20, (50, 2)).astype(np.float32)
    pos[0] = (2.5, -3.5)  # halfway cases round away from zero after the canvas offset
    roi, sub = orc.mls_roi(pos, m, rows, cols)
    cx = np.float32(np.ceil(rows / 2 - m / 2))
    cy = np.float32(np.ceil(cols / 2 - m / 2))
    p1 = pos[:, 0] + cx
    p2 = pos[:, 1] + cy
    # origin + 1 + fraction reproduces the float32 position exactly; fraction in [-1/2, 1/2]
    assert np.array_equal((roi[:, 0] + 1).astype(np.float32) + sub[:, 1], p1)
    assert np.array_equal((roi[:, 1] + 1).astype(np.float32) + sub[:, 0], p2)
    assert np.all(np.abs(sub) <= 0.5)
    assert roi[0, 0] + 1 == int(np.floor(p1[0] + 0.5)) and sub[0, 1] == -0.5
    assert roi[0, 1] + 1 == int(np.floor(p2[0] + 0.5))


def test_init_normalisation_closed_forms(orc):
    P = make_mls_problem(k=3)
    st = P["state"]
    m = P["m"]
    # illumination = sum of |P0|^2 placed at every ROI, rescaled with the probe
    ill = np.zeros(P["shape"])
    for (r0, c0) in st.roi:
        ill[r0:r0 + m, c0:c0 + m] += np.abs(st.probe[0]) ** 2
    assert np.allclose(st.illum, ill, rtol=1e-12)
    assert np.isclose(st.max_illum, ill.max())
    # object: illumination-weighted mean |O|^2 = 1
    assert np.isclose(np.sum(st.illum * np.abs(st.obj) ** 2) / np.sum(st.illum), 1.0)
    # probe power: median over frames of M^2 sum|P|^2 / sum I = 1
    tot = np.sum(np.abs(st.probe) ** 2)
    assert np.isclose(np.median(m * m * tot / P["frame_power"]), 1.0)
    # variable mode unit rms before the probe scaling; evolution rescaled by the same factor
    v0 = P["vmode0"]
    rv = np.sqrt(np.mean(np.abs(v0) ** 2))
    corr = np.sqrt(np.sum(np.abs(P["probe0"]) ** 2) / tot)
    assert np.allclose(st.vmode, v0 / rv / corr)
    assert np.allclose(st.evo[1], P["evo1_0"] * rv)
    assert np.all(st.evo[0] == 1.0)


@pytest.mark.parametrize("it", [1, 3])
def test_consistent_data_is_a_fixed_point(orc, it):
    """A = |FFT(model)|: chi ~ 0, so no group changes the model's exit waves (calc_object_norm
    moves a scale between O and P, which the exit waves do not see)."""
    P = make_mls_problem(k=2)
    st = P["state"]
    before = mls_exit_amplitudes(st, orc)
    prm = orc.mls_params(P["box"], iter=it, pps=it)
    st.iteration(prm, mls_groups(P["n"], 4, np.random.default_rng(0)))
    after = mls_exit_amplitudes(st, orc)
    assert np.max(np.abs(after - before)) < 1e-6 * np.max(before)
    assert np.max(np.abs(st.pos - P["pos"])) < 1e-5


def test_single_frame_updates_closed_form(orc):
    """One frame, one mode, no sub-pixel shift, no variable probe, beta_LSQ = 0 (steps stay 1):
    O += (conj(P) chi + i eps) / (p + delta max p),  P += p/(p + 0.01 max p) chi conj(O),
    chi = Pi(P O) - P O with Pi the Fourier modulus projection, p = |P|^2 after the running
    average (the initial p is |P|^2 too)."""
    rng = np.random.default_rng(8)
    m = 16
    rows = cols = 24
    obj = np.exp(1j * rng.uniform(-1, 1, (rows, cols)))
    probe = (np.exp(-((np.mgrid[0:m, 0:m] - m / 2) ** 2).sum(0) / 20.0) * (1 + 0.2j))[None]
    pos = np.zeros((1, 2), np.float32)
    st = orc.MlsState(obj, probe, np.ones((m, m)), np.zeros(1), pos, np.zeros((1, m, m)))
    st.init(np.array([np.sum(np.abs(probe) ** 2) * m * m]), variable_probe_modes=0)
    amp = np.abs(np.fft.fft2(_field(rng, m)))  # inconsistent data
    st.amp = amp[None].copy()
    r0, c0 = st.roi[0]
    O0, P0 = st.obj.copy(), st.probe[0].copy()
    box = (r0 + 2, r0 + 9, c0 + 2, c0 + 9)  # |O| = 1 there: calc_object_norm is the identity
    prm = orc.mls_params(box, iter=1, pps=99, variable_probe=0, variable_intensity=0, variable_probe_modes=0,
                         apply_subpix_shift=0, beta_lsq=0.0)
    st.iteration(prm, [np.array([0])])
    op = O0[r0:r0 + m, c0:c0 + m]
    psi = P0 * op
    F = np.fft.fft2(psi)
    chi = np.fft.ifft2(amp * F / np.abs(F)) - psi
    p = np.abs(P0) ** 2
    mx = p.max()
    eps = float(np.float32(np.finfo(np.float64).eps))
    o_exp = op + (np.conj(P0) * chi + 1j * eps) / (p + 0.1 * mx)
    p_exp = P0 + p / (p + 0.01 * mx) * chi * np.conj(op)
    assert np.allclose(st.obj[r0:r0 + m, c0:c0 + m], o_exp, atol=1e-10)
    assert np.allclose(st.probe[0], p_exp, atol=1e-10)
    # outside the ROI only the i*eps/(delta max p) constant is added
    out = np.ones((rows, cols), bool)
    out[r0:r0 + m, c0:c0 + m] = False
    assert np.allclose(st.obj[out], O0[out] + 1j * eps / (0.1 * mx), atol=1e-15)
