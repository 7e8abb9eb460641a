"""Pins of the FP64 network oracle (oracle/oracle_nn.cpp): one convolution against scipy's
correlation, the full backward pass against central finite differences."""
import numpy as np
from scipy import signal


def test_conv2d_matches_scipy_correlation(orc):
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 3, 9, 9))
    w = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    y = orc.nn_conv2d(x, w, b, relu=False)
    for n in range(2):
        for co in range(4):
            ref = sum(signal.correlate2d(x[n, ci], w[co, ci], mode="same") for ci in range(3)) + b[co]
            assert np.allclose(y[n, co], ref, atol=1e-12)
    yr = orc.nn_conv2d(x, w, b, relu=True)
    assert np.allclose(yr, np.maximum(y, 0))


def test_backward_matches_finite_differences(orc):
    rng = np.random.default_rng(1)
    c, s, batch = 2, 16, 2
    n = orc.nn_num_params(c)
    p = (rng.standard_normal(n) * 0.3).astype(np.float32)
    x = rng.standard_normal((batch, s, s)).astype(np.float32)
    t = rng.standard_normal((batch, 2, s, s)).astype(np.float32)
    _, loss, g = orc.nn_forward_backward(p, x, t, c)
    idx = rng.choice(n, 40, replace=False)
    h = 1e-5  # small enough that no ReLU or max-pool switches inside the step
    for i in idx:
        pp, pm = p.astype(np.float64).copy(), p.astype(np.float64).copy()
        # float32 parameters: step on a float32 grid point to keep the perturbation exact
        pp[i] += h
        pm[i] -= h
        lp = orc.nn_forward_backward(pp.astype(np.float32), x, t, c, grad=False)[1]
        lm = orc.nn_forward_backward(pm.astype(np.float32), x, t, c, grad=False)[1]
        hp = float(np.float32(pp[i])) - float(np.float32(p[i]))
        hm = float(np.float32(p[i])) - float(np.float32(pm[i]))
        fd = (lp - lm) / (hp + hm)
        assert abs(fd - g[i]) <= 1e-4 * max(1e-3, abs(g[i])) + 1e-7, (i, fd, g[i])


def test_weighted_loss_closed_form_and_gradient(orc):
    """An indicator weight gives the plain mean squared error over the marked pixels, and the
    weighted gradient agrees with finite differences of the weighted loss."""
    rng = np.random.default_rng(2)
    c, s, batch = 2, 16, 2
    n = orc.nn_num_params(c)
    p = (rng.standard_normal(n) * 0.3).astype(np.float32)
    x = rng.standard_normal((batch, s, s)).astype(np.float32)
    t = rng.standard_normal((batch, 2, s, s)).astype(np.float32)
    mask = np.zeros((s, s), np.float32)
    mask[3:11, 5:9] = 1.0
    out, loss, _ = orc.nn_forward_backward(p, x, t, c, grad=False, weight=mask)
    d2 = (out - t) ** 2
    assert np.isclose(loss, d2[:, :, mask > 0].mean(), rtol=1e-12)
    w = rng.uniform(0, 2, (s, s)).astype(np.float32)
    _, lw, g = orc.nn_forward_backward(p, x, t, c, weight=w)
    assert np.isclose(lw, (d2 * w).sum() / (batch * 2 * w.astype(np.float64).sum()), rtol=1e-12)
    for i in rng.choice(n, 12, replace=False):
        pp, pm = p.astype(np.float64).copy(), p.astype(np.float64).copy()
        pp[i] += 1e-5
        pm[i] -= 1e-5
        lp = orc.nn_forward_backward(pp.astype(np.float32), x, t, c, grad=False, weight=w)[1]
        lm = orc.nn_forward_backward(pm.astype(np.float32), x, t, c, grad=False, weight=w)[1]
        fd = (lp - lm) / (float(np.float32(pp[i])) - float(np.float32(pm[i])))
        assert abs(fd - g[i]) <= 1e-4 * max(1e-3, abs(g[i])) + 1e-7, (i, fd, g[i])
