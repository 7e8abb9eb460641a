"""Learned warm start for MLs: network binding, training data, stitching.

The network (include/ptyk.h, ``ptyk_net_*``) maps a log-scaled diffraction amplitude to the
amplitude and phase of the object patch under that frame; ``warm_start`` runs it on the GPU
(preprocessing included) and has the engine average the predicted patches into its object
(``ptyk_stitch_patches``), weighting each by the probe intensity.  The plain iterative solver
then refines that object, and remains the fallback when no network is given.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _engine, sim


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Net:
    """Encoder-decoder on the GPU: size x size inputs, base channel width c, fixed batch."""

    def __init__(self, size: int = 256, c: int = 32, batch: int = 8, seed: int = 0, device: int = 0):
        self.L = _engine.load()
        self.size, self.c, self.batch = int(size), int(c), int(batch)
        h = C.c_void_p()
        rc = self.L.ptyk_net_create(device, self.size, self.c, self.batch, seed, C.byref(h))
        if rc != 0:
            raise _engine.PtykError(rc, self.L.ptyk_net_last_error(None).decode())
        self.h = h
        self.n_params = int(self.L.ptyk_net_num_params(h))
        self.n_data = 0

    def _chk(self, rc):
        if rc != 0:
            raise _engine.PtykError(rc, self.L.ptyk_net_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.ptyk_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def get_params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        self._chk(self.L.ptyk_net_get_params(self.h, _ptr(out)))
        return out

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        assert a.size == self.n_params
        self._chk(self.L.ptyk_net_set_params(self.h, _ptr(a)))

    def set_data(self, inputs, targets):
        x = np.ascontiguousarray(inputs, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        n = x.shape[0]
        assert x.shape == (n, self.size, self.size) and t.shape == (n, 2, self.size, self.size)
        self._chk(self.L.ptyk_net_set_data(self.h, _ptr(x), _ptr(t), n))
        self.n_data = n

    def set_loss_weight(self, weight=None):
        """S x S per-pixel loss weight shared by both outputs (None: uniform)."""
        if weight is None:
            self._chk(self.L.ptyk_net_set_loss_weight(self.h, None))
            return
        w = np.ascontiguousarray(weight, dtype=np.float32)
        assert w.shape == (self.size, self.size)
        self._chk(self.L.ptyk_net_set_loss_weight(self.h, _ptr(w)))

    def reserve_data(self, n: int):
        self._chk(self.L.ptyk_net_reserve_data(self.h, int(n)))
        self.n_data = 0
        self.data_cap = int(n)

    def put_data(self, offset: int, inputs, targets):
        x = np.ascontiguousarray(inputs, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        n = x.shape[0]
        assert x.shape == (n, self.size, self.size) and t.shape == (n, 2, self.size, self.size)
        self._chk(self.L.ptyk_net_put_data(self.h, int(offset), _ptr(x), _ptr(t), n))
        self.n_data = max(self.n_data, int(offset) + n)

    def train(self, order, lr: float) -> float:
        o = np.ascontiguousarray(order, dtype=np.int64)
        loss = C.c_double(0.0)
        self._chk(self.L.ptyk_net_train(self.h, _ptr(o), o.size, float(lr), C.byref(loss)))
        return loss.value

    def evaluate(self, order) -> float:
        o = np.ascontiguousarray(order, dtype=np.int64)
        loss = C.c_double(0.0)
        self._chk(self.L.ptyk_net_evaluate(self.h, _ptr(o), o.size, C.byref(loss)))
        return loss.value

    def predict(self, inputs) -> np.ndarray:
        x = np.ascontiguousarray(inputs, dtype=np.float32)
        out = np.empty((x.shape[0], 2, self.size, self.size), np.float32)
        self._chk(self.L.ptyk_net_predict(self.h, _ptr(x), x.shape[0], _ptr(out)))
        return out

    def predict_amplitudes(self, amplitudes) -> np.ndarray:
        """predict(preprocess(amplitudes)) with the preprocessing done on the GPU."""
        a = np.ascontiguousarray(amplitudes, dtype=np.float32)
        out = np.empty((a.shape[0], 2, self.size, self.size), np.float32)
        self._chk(self.L.ptyk_net_predict_amplitudes(self.h, _ptr(a), a.shape[0], _ptr(out)))
        return out

    def gradient(self, inputs, targets):
        x = np.ascontiguousarray(inputs, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        assert x.shape[0] == self.batch and t.shape[0] == self.batch
        loss = C.c_double(0.0)
        g = np.empty(self.n_params, np.float32)
        self._chk(self.L.ptyk_net_gradient(self.h, _ptr(x), _ptr(t), C.byref(loss), _ptr(g)))
        return loss.value, g


def net_forward_flops(s: int, c: int) -> float:
    """Multiply-add FLOPs of one forward pass (2 x C_in x C_out x taps per output pixel, summed over
    the 15 convolutions); a training step costs about three times this."""
    tab = [(1, c, 3, 0), (c, c, 3, 0), (c, 2 * c, 3, 1), (2 * c, 2 * c, 3, 1), (2 * c, 4 * c, 3, 2),
           (4 * c, 4 * c, 3, 2), (4 * c, 8 * c, 3, 3), (8 * c, 8 * c, 3, 3), (12 * c, 4 * c, 3, 2),
           (4 * c, 4 * c, 3, 2), (6 * c, 2 * c, 3, 1), (2 * c, 2 * c, 3, 1), (3 * c, c, 3, 0), (c, c, 3, 0),
           (c, 2, 1, 0)]
    return float(sum(2 * ci * co * k * k * (s >> lv) ** 2 for ci, co, k, lv in tab))


def preprocess(amp) -> np.ndarray:
    """log(1 + A) scaled per frame to a maximum of 1 (the network's input)."""
    a = np.log1p(np.maximum(np.asarray(amp, np.float32), 0))
    mx = a.reshape(a.shape[0], -1).max(axis=1)
    return (a / np.maximum(mx, 1e-12)[:, None, None]).astype(np.float32)


def patch_targets(obj, roi, m: int) -> np.ndarray:
    """(n, 2, m, m): amplitude and phase of the object under each ROI (integer origins)."""
    out = np.empty((len(roi), 2, m, m), np.float32)
    for j, (r0, c0) in enumerate(roi):
        p = obj[r0:r0 + m, c0:c0 + m]
        out[j, 0] = np.abs(p)
        out[j, 1] = np.angle(p)
    return out


def _simulated_chunk(e, modes, side, cnt, photons, rng, phase_std, phase_corr=6.0):
    k, m, _ = modes.shape
    obj = sim.random_field_object((side, side), rng, phase_std=float(rng.uniform(*phase_std)),
                                  phase_corr=phase_corr)
    roi = rng.integers(0, side - m + 1, size=(cnt, 2))
    inten = sum(e.simulate(obj, modes[i], roi.astype(np.float32), squared=True) for i in range(k))
    amp = sim.poisson_amplitudes(inten, photons, rng)
    return preprocess(amp), patch_targets(obj, roi, m)


def training_set(net: Net, n_frames: int, modes, photons: float = 1e5, frames_per_object: int = 200,
                 seed: int = 0, phase_std=(0.2, 1.0), device: int = 0, log=None):
    """Fill the network's sample store with n_frames frames that the engine's simulator forms
    from random-field objects (phase std drawn from phase_std per object) under the incoherent
    probe modes (K x M x M), Poisson-sampled at `photons` per frame.  Sample j's input is the
    preprocessed amplitude, its target the object patch's (amplitude, phase)."""
    k, m, _ = modes.shape
    assert m == net.size
    rng = np.random.default_rng(seed)
    net.reserve_data(n_frames)
    side = 4 * m
    with _engine.Engine(1, m, (m, m), device=device) as e:
        done = 0
        while done < n_frames:
            cnt = min(frames_per_object, n_frames - done)
            x, t = _simulated_chunk(e, modes, side, cnt, photons, rng, phase_std)
            net.put_data(done, x, t)
            done += cnt
            if log and (done // frames_per_object) % 25 == 0:
                log(f"simulated {done}/{n_frames} frames")


def fit(net: Net, n_train: int, epochs: int, lr: float = 1e-3, seed: int = 0, held_out: int = 0, log=None):
    """Adam over samples [0, n_train) for `epochs` shuffled passes; held-out loss over
    [n_train, n_train + held_out) after each epoch.  Returns [(train_loss, held_out_loss)]."""
    rng = np.random.default_rng(seed)
    b = net.batch
    n_use = n_train // b * b
    val = np.arange(n_train, n_train + held_out // b * b)
    hist = []
    for ep in range(epochs):
        tl = net.train(rng.permutation(n_train)[:n_use], lr)
        vl = net.evaluate(val) if len(val) else float("nan")
        hist.append((tl, vl))
        if log:
            log(f"epoch {ep + 1}/{epochs}: train {tl:.5f} held-out {vl:.5f}")
    return hist


def lattice_frames(n_rows: int, n_cols: int, stride: int) -> np.ndarray:
    """Frame ids of a row-major n_rows x n_cols raster on every stride-th row and column: a sparser
    set of patches that still covers the scan when stride x step is below the beam width."""
    r, c = np.meshgrid(np.arange(0, n_rows, stride), np.arange(0, n_cols, stride), indexing="ij")
    return (r * n_cols + c).ravel().astype(np.int32)


def warm_start(engine: _engine.Engine, net: Net, amplitudes, frames, rel_threshold: float = 1e-3, fill=1.0 + 0j):
    """Initial object for the engine's scan: the network's patches for the given frames (from the
    raw amplitudes, preprocessed on the GPU), stitched on the GPU at the frames' ROIs and weighted
    by the current probe intensity.  Call after positions (or ptyk_mls_begin) are set."""
    fr = np.ascontiguousarray(frames, dtype=np.int32)
    pred = net.predict_amplitudes(np.asarray(amplitudes)[fr])
    engine.stitch_patches(pred, fr, rel_threshold, fill)
    return pred


def _mls_to_error(e: _engine.Engine, W: dict, vmode, evo1, fp, target: float, max_iter: int, seed: int,
                  warm=None):
    """MLs from the engine's current object until e <= target or max_iter; wall time of the
    solver (and, for a warm start, of inference and stitching, run inside the clock)."""
    import time
    n = W["n"]
    e.set_probe(W["probe_true"])
    e.mls_begin(_engine.mls_config(W["box"]), W["pos"], vmode, evo1, fp)
    t = 0.0
    if warm is not None:
        net, amp, frames = warm
        t0 = time.perf_counter()
        warm_start(e, net, amp, frames)
        e.set_probe(W["probe_true"])
        e.mls_begin(_engine.mls_config(W["box"]), W["pos"], vmode, evo1, fp)
        e.synchronize()
        t += time.perf_counter() - t0
    from . import _driver
    parts = _driver.mls_partitions(n, max_iter, seed)
    rng = np.random.default_rng(seed)
    hist = [(0, 0.0, e.object_error()[0])]
    reached = None
    for it in range(1, max_iter + 1):
        p = parts[it % len(parts)]
        t0 = time.perf_counter()
        e.mls_iteration(it, [p[j] for j in rng.permutation(len(p))])
        e.synchronize()
        t += time.perf_counter() - t0
        err = e.object_error()[0]
        hist.append((it, t, err))
        if reached is None and err <= target:
            reached = (it, t)
            break
    return reached, hist


def c4_experiment(n_train: int = 50000, n_held: int = 5000, epochs: int = 2, lr: float = 3e-4, batch: int = 16,
                  c: int = 32, target: float = 1e-2, max_iter: int = 100, stride: int = 2, seed: int = 0,
                  log=None) -> dict:
    """Learned warm start against a plain uniform start on the 51 x 51 / 256^2 / 5-mode MLs
    workload: train on simulated random-field frames (phase std 0.2-1.0 rad) under the workload's
    probe modes with a beam-weighted loss, report the held-out loss beside the constant
    predictor's, then run MLs from both starts to the target normalised object error."""
    from . import workloads
    W = workloads.mls_c2()
    n, m, k = W["n"], W["m"], W["k"]
    modes = W["probe_true"]
    pint = (np.abs(modes.astype(np.complex128)) ** 2).sum(axis=0)
    wgt = (pint / pint.max()).astype(np.float32)
    net = Net(size=m, c=c, batch=batch, seed=seed + 1)
    net.set_loss_weight(wgt)
    # training and held-out frames, streamed into the network's sample store; the held-out
    # targets' weighted moments give the constant predictor's loss
    rng = np.random.default_rng(seed + 5)
    net.reserve_data(n_train + n_held)
    s1 = np.zeros(2)
    s2 = np.zeros(2)
    swt = 0.0
    with _engine.Engine(1, m, (m, m)) as es:
        done = 0
        while done < n_train + n_held:
            cnt = min(200, n_train + n_held - done)
            x, t = _simulated_chunk(es, modes, 4 * m, cnt, W["photons"], rng, (0.2, 1.0))
            net.put_data(done, x, t)
            lo = max(done, n_train) - done
            if lo < cnt:
                tt = t[lo:].astype(np.float64)
                s1 += (tt * wgt).sum(axis=(0, 2, 3))
                s2 += (tt * tt * wgt).sum(axis=(0, 2, 3))
                swt += float(wgt.sum()) * (cnt - lo)
            done += cnt
            if log and done % 5000 < cnt:
                log(f"simulated {done}/{n_train + n_held}")
    const = float(((s2 - s1 * s1 / swt) / (2 * swt)).sum())
    import time
    t0 = time.perf_counter()
    hist = fit(net, n_train, epochs, lr, seed=seed, held_out=n_held, log=log)
    t_fit = time.perf_counter() - t0
    e = _engine.Engine(n, m, W["canvas"], n_modes=k)
    e.set_frames(W["amp"])
    amp = e.get_frames()
    e.set_true_object(W["obj_true"])
    r = np.random.default_rng(seed)
    vmode = (r.uniform(0, 1, (m, m)) * np.exp(2j * np.pi * r.uniform(0, 1, (m, m)))).astype(np.complex64)
    evo1 = 1e-6 * r.standard_normal(n)
    fp = (amp.astype(np.float64) ** 2).sum(axis=(1, 2))
    side = int(round(np.sqrt(n)))
    frames = lattice_frames(side, side, stride)
    e.set_object(np.ones(W["canvas"], np.complex64))
    plain, plain_hist = _mls_to_error(e, W, vmode, evo1, fp, target, max_iter, seed)
    warm, warm_hist = _mls_to_error(e, W, vmode, evo1, fp, target, max_iter, seed, warm=(net, amp, frames))
    e.close()
    net.close()
    best = lambda h: min(v for _, _, v in h)
    gf = net_forward_flops(m, c)
    train_rate = epochs * (n_train // batch * batch) / t_fit  # includes the held-out passes
    return {
        "train_frames": n_train, "held_out_frames": n_held, "epochs": epochs,
        "train_samples_per_s": train_rate, "train_tflops": 3 * gf * train_rate / 1e12,
        "forward_gflop_per_sample": gf / 1e9,
        "held_out_loss": hist[-1][1], "constant_predictor_loss": const, "train_loss": hist[-1][0],
        "target_error": target, "max_iterations": max_iter, "inference_frames": int(len(frames)),
        "plain": {"reached": plain, "e0": plain_hist[0][2], "best": best(plain_hist),
                  "trace": plain_hist[::5]},
        "warm": {"reached": warm, "e0": warm_hist[0][2], "best": best(warm_hist), "trace": warm_hist[::5]},
        "speedup": (plain[1] / warm[1]) if (plain and warm) else None,
    }
