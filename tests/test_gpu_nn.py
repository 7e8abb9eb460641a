"""Warm-start network on the GPU against the FP64 oracle (BF16 operand rounding emulated).

The GPU rounds weights, inputs, targets and stored activations to BF16 and accumulates in FP32;
the oracle rounds the same operands and accumulates in FP64, so the two differ only by
accumulation order and the BF16 rounding flips that order causes.  Those flips propagate through
the 15 layers: on the oracle alone, perturbing the weights by 1e-6 (relative) moves the emulated
output by 4e-3 to 9e-3 and single layers' gradients by up to 6e-2 (relative L2) at these sizes.
So each gradient test measures that noise floor for its own inputs (one oracle run with weights
perturbed by 1e-6) and allows 3x it: output <= max(1.5e-2, 3 x floor), loss <= 1e-2 relative,
each layer's weight and bias gradient <= max(2e-2, 3 x that layer's floor).  A wrong index or a
missing term gives O(1).
"""
import numpy as np
import pytest

import oracle
from ptychopy_amd import warmstart

pytestmark = pytest.mark.gpu


def layer_slices(c):
    tab = [(1, c, 3), (c, c, 3), (c, 2 * c, 3), (2 * c, 2 * c, 3), (2 * c, 4 * c, 3), (4 * c, 4 * c, 3),
           (4 * c, 8 * c, 3), (8 * c, 8 * c, 3), (12 * c, 4 * c, 3), (4 * c, 4 * c, 3), (6 * c, 2 * c, 3),
           (2 * c, 2 * c, 3), (3 * c, c, 3), (c, c, 3), (c, 2, 1)]
    off, out = 0, []
    for cin, cout, ks in tab:
        w = slice(off, off + cout * cin * ks * ks)
        off = w.stop
        b = slice(off, off + cout)
        off = b.stop
        out.append((w, b))
    return out, off


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def to_bf16(a):
    """Round-to-nearest-even to BF16, returned as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32)


def sample(rng, batch, s):
    x = rng.uniform(0, 1, size=(batch, s, s)).astype(np.float32)
    t = np.stack([rng.uniform(0.5, 1.0, size=(batch, s, s)), rng.uniform(-1, 1, size=(batch, s, s))],
                 axis=1).astype(np.float32)
    return x, t


@pytest.mark.parametrize("s,c,batch,weighted", [(16, 4, 2, False), (32, 8, 3, False), (64, 16, 2, False),
                                                (32, 8, 2, True), (256, 8, 2, True), (32, 32, 2, True)])
def test_gradient_matches_oracle(s, c, batch, weighted):
    print(f"net gradient s={s} c={c} batch={batch} weighted={weighted}", flush=True)
    rng = np.random.default_rng(7 + s)
    x, t = sample(rng, batch, s)
    w = None
    if weighted:  # a Gaussian beam profile, as the warm start uses
        yy, xx = np.mgrid[0:s, 0:s] - s / 2
        w = np.exp(-(xx ** 2 + yy ** 2) / (2 * (s / 6) ** 2)).astype(np.float32)
    with warmstart.Net(size=s, c=c, batch=batch, seed=3) as net:
        net.set_loss_weight(w)
        sl, n = layer_slices(c)
        assert net.n_params == n == oracle.nn_num_params(c)
        p = net.get_params()
        # biases start at zero; give them values so their path is exercised
        for _, b in sl:
            p[b] = rng.normal(0, 0.05, size=b.stop - b.start)
        net.set_params(p)
        loss, g = net.gradient(x, t)
        out = net.predict(x)
    ref_out, ref_loss, ref_g = oracle.nn_forward_backward(p, x, t, c, emulate_bf16=True, weight=w)
    # the emulation's own sensitivity to 1e-6 relative changes (accumulation-order scale)
    pp = (p * (1 + 1e-6 * np.random.default_rng(99).standard_normal(p.size))).astype(np.float32)
    n_out, _, n_g = oracle.nn_forward_backward(pp, x, t, c, emulate_bf16=True, weight=w)
    floor_out = rel(n_out, ref_out)
    print(f"output: GPU {rel(out, ref_out):.2e}, noise floor {floor_out:.2e}")
    assert rel(out, ref_out) <= max(1.5e-2, 3 * floor_out)
    assert abs(loss - ref_loss) <= 1e-2 * ref_loss
    for li, (wl, bl) in enumerate(sl):  # each layer against its own noise floor
        for part in (wl, bl):
            err, floor = rel(g[part], ref_g[part]), rel(n_g[part], ref_g[part])
            assert err <= max(2e-2, 3 * floor), (li, err, floor)
    # the unrounded FP64 network is close too (BF16 is the only approximation)
    f_out, f_loss, _ = oracle.nn_forward_backward(p, x, t, c, emulate_bf16=False)
    assert rel(out, f_out) <= 5e-2


def test_predict_handles_a_ragged_last_batch():
    rng = np.random.default_rng(1)
    s, c, batch = 16, 4, 4
    x, _ = sample(rng, 6, s)
    with warmstart.Net(size=s, c=c, batch=batch, seed=5) as net:
        p = net.get_params()
        y6 = net.predict(x)
        y2 = net.predict(x[4:])
    np.testing.assert_array_equal(y6[4:], y2)  # samples do not interact inside a batch
    ref, _, _ = oracle.nn_forward_backward(p, x, np.zeros((6, 2, s, s), np.float32), c, emulate_bf16=True,
                                           grad=False)
    assert rel(y6, ref) <= 1.5e-2


def test_first_adam_step_is_a_signed_learning_rate_step():
    """At step 1 Adam's bias-corrected moments are g and g^2, so w1 = w0 - lr g / (|g| + 1e-8)."""
    rng = np.random.default_rng(2)
    s, c, batch = 16, 4, 2
    x, t = sample(rng, batch, s)
    lr = 1e-3
    with warmstart.Net(size=s, c=c, batch=batch, seed=9) as net:
        w0 = net.get_params()
        _, g = net.gradient(x, t)
        net.set_data(x, t)
        net.train(np.arange(batch), lr)
        w1 = net.get_params()
    expect = w0 - lr * g / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(w1, expect, rtol=0, atol=1e-6)


def test_training_reduces_the_loss_and_evaluate_agrees():
    rng = np.random.default_rng(3)
    s, c, batch, n = 32, 8, 4, 64
    xs, ts = [], []
    for _ in range(n):  # a learnable map: amplitude 1 - 0.3 x, phase x - 0.5
        x = rng.uniform(0, 1, size=(s, s)).astype(np.float32)
        xs.append(x)
        ts.append(np.stack([1 - 0.3 * x, x - 0.5]))
    xs, ts = np.array(xs, np.float32), np.array(ts, np.float32)
    with warmstart.Net(size=s, c=c, batch=batch, seed=11) as net:
        net.set_data(xs, ts)
        order = np.arange(n)
        first = net.evaluate(order)
        for _ in range(15):
            net.train(rng.permutation(n), 2e-3)
        last = net.evaluate(order)
        head = net.evaluate(np.arange(batch))
        pred = net.predict(xs[:batch])
    assert last < 0.25 * first, (first, last)
    # evaluate() scores against the BF16 copy of the targets; predict() gives the same outputs
    direct = float(np.mean((pred.astype(np.float64) - to_bf16(ts[:batch])) ** 2))
    assert abs(head - direct) <= 1e-6 + 1e-5 * direct


def test_invalid_arguments_are_rejected():
    with pytest.raises(Exception, match="power of two"):
        warmstart.Net(size=24, c=4, batch=2)
    with warmstart.Net(size=16, c=4, batch=2) as net:
        x, t = sample(np.random.default_rng(0), 2, 16)
        net.set_data(x, t)
        with pytest.raises(Exception, match="multiple of the batch"):
            net.train(np.arange(3), 1e-3)
        with pytest.raises(Exception, match="out of range"):
            net.train(np.array([0, 5]), 1e-3)
        with pytest.raises(Exception, match="non-negative"):
            net.set_loss_weight(-np.ones((16, 16), np.float32))
        with pytest.raises(Exception, match="all zero"):
            net.set_loss_weight(np.zeros((16, 16), np.float32))
        with pytest.raises(Exception, match="beyond the reserved"):
            net.put_data(1, x, t)


def test_streamed_samples_equal_one_shot_samples():
    rng = np.random.default_rng(4)
    s, c, batch = 16, 4, 2
    x, t = sample(rng, 6, s)
    with warmstart.Net(size=s, c=c, batch=batch, seed=5) as a, warmstart.Net(size=s, c=c, batch=batch, seed=5) as b:
        a.set_data(x, t)
        b.reserve_data(6)
        b.put_data(4, x[4:], t[4:])
        b.put_data(0, x[:4], t[:4])
        order = np.array([5, 0, 3, 2, 1, 4])
        assert a.train(order, 1e-3) == b.train(order, 1e-3)
        np.testing.assert_array_equal(a.get_params(), b.get_params())
