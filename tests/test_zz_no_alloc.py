"""ptyk_simulate_noalloc (device inputs and outputs, exit waves staged in the library's fixed
device array) against ptyk_simulate on the same inputs, through the C ABI: two calls in a row, the
second with fewer frames, so the second call runs over what the first left in the array; and a
request larger than the array is refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(n, m, rows, cols, seed):
    rng = np.random.default_rng(seed)
    obj = (rng.uniform(0.5, 1.0, (rows, cols)) * np.exp(1j * rng.uniform(-1, 1, (rows, cols)))).astype(np.complex64)
    yy, xx = np.mgrid[0:m, 0:m] - m / 2
    probe = (np.exp(-(xx ** 2 + yy ** 2) / (2 * (m / 6) ** 2)) * np.exp(0.3j * xx / m)).astype(np.complex64)
    pos = np.stack([rng.uniform(0, rows - m, n), rng.uniform(0, cols - m, n)], axis=1).astype(np.float32)
    return obj, probe, pos


def _noalloc(e, obj, probe, pos, squared):
    import torch
    dev = torch.device("cuda", 0)
    o = torch.from_numpy(obj.view(np.float32).copy()).to(dev)
    p = torch.from_numpy(probe.view(np.float32).copy()).to(dev)
    ps = torch.from_numpy(pos.copy()).to(dev)
    out = torch.empty((pos.shape[0], e.m, e.m), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    v = C.c_void_p
    rc = e.lib.ptyk_simulate_noalloc(e.h, v(o.data_ptr()), obj.shape[0], obj.shape[1], v(p.data_ptr()),
                                     v(ps.data_ptr()), pos.shape[0], int(squared), v(out.data_ptr()))
    assert rc == 0, e.lib.ptyk_last_error(e.h)
    assert e.lib.ptyk_synchronize(e.h) == 0
    return out.cpu().numpy()


@pytest.mark.parametrize("m,squared", [(64, 1), (256, 0)])
def test_noalloc_simulate_matches_simulate(gpu, m, squared):
    from ptychopy_amd._engine import Engine
    rows, cols = m + 40, m + 56
    with Engine(1, m, (m, m)) as e:
        for call, n in enumerate((13, 5)):
            obj, probe, pos = _case(n, m, rows, cols, seed=100 * m + call)
            print(f"no-alloc simulate m={m} call={call} n={n}", flush=True)
            ref = e.simulate(obj, probe, pos, squared=bool(squared))
            got = _noalloc(e, obj, probe, pos, squared)
            rel = float(np.abs(got - ref).max() / np.abs(ref).max())
            print(f"  max relative difference {rel:.3g}, bitwise {np.array_equal(got, ref)}", flush=True)
            assert rel <= 1e-6


def test_noalloc_simulate_refuses_more_than_the_array_holds(gpu):
    from ptychopy_amd._engine import Engine
    m = 512
    cap = 64 << 20
    n = cap // (m * m * 8) + 1  # one frame more than fits
    with Engine(1, m, (m, m)) as e:
        v = C.c_void_p
        dummy = C.c_void_p(16)  # never dereferenced: the size check runs first
        rc = e.lib.ptyk_simulate_noalloc(e.h, dummy, m, m, dummy, dummy, n, 0, dummy)
        assert rc == -1
        assert b"scratch" in e.lib.ptyk_last_error(e.h)
