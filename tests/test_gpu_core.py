"""GPU parity of the engine's building blocks against the FP64 oracle (through the C ABI)."""
import numpy as np
import pytest

from conftest import rel_err
from problems import make_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m", [16, 32, 64, 128, 256, 512])
def test_fft2_matches_oracle(gpu, orc, m):
    from ptychopy_amd._engine import Engine
    rng = np.random.default_rng(m)
    n = 5 if m <= 128 else 2
    x = (rng.standard_normal((n, m, m)) + 1j * rng.standard_normal((n, m, m))).astype(np.complex64)
    with Engine(1, m, (m, m)) as e:
        f = e.fft2(x)
        assert rel_err(f, orc.fft2(x)) < 2e-6
        b = e.fft2(x, inverse=True)
        assert rel_err(b, orc.fft2(x, inverse=True)) < 2e-6


def test_fft2_tone_and_delta(gpu):
    from ptychopy_amd._engine import Engine
    m = 128
    i, j = np.mgrid[0:m, 0:m]
    tone = np.exp(2j * np.pi * (3 * i + 17 * j) / m).astype(np.complex64)
    with Engine(1, m, (m, m)) as e:
        f = e.fft2(tone)
    expect = np.zeros((m, m), complex)
    expect[3, 17] = m * m
    assert np.abs(f - expect).max() < 1e-2  # FP32: 16384 * ~1e-6


def test_frame_codec_roundtrip(gpu):
    from ptychopy_amd._engine import Engine
    rng = np.random.default_rng(0)
    n, m = 4, 32
    a = (rng.uniform(0, 1, (n, m, m)) ** 4 * 10.0 ** rng.uniform(-2, 4, (n, 1, 1))).astype(np.float32)
    a[1] = 0.0  # an all-zero frame keeps scale 1 and decodes to zeros
    a[2, 0, 0] = 0.0
    with Engine(n, m, (m, m)) as e:
        e.set_frames(a)
        d = e.get_frames()
        sc = e.get_frame_scales()
    assert np.all(np.log2(sc) == np.round(np.log2(sc)))  # power-of-two scales
    assert np.all(d[1] == 0) and d[2, 0, 0] == 0.0
    mx = a.reshape(n, -1).max(axis=1)
    for f in (0, 2, 3):
        big = a[f] > mx[f] * 2 ** -14  # normal FP16 range after scaling
        assert np.all(np.abs(d[f][big] - a[f][big]) <= 2 ** -11 * a[f][big] + 1e-30)
        assert np.max(np.abs(d[f] - a[f])) <= mx[f] * 2 ** -11


def test_simulate_matches_oracle(gpu, orc):
    from ptychopy_amd._engine import Engine
    p = make_problem(m=64, n_side=3, seed=1)
    with Engine(p["n"], p["m"], p["shape"]) as e:
        g = e.simulate(p["obj"], p["probe"], p["pos"], squared=True)
    ref = orc.simulate(p["obj"], p["probe"], p["pos"], squared=True)
    assert rel_err(g, ref) < 1e-5


def test_probe_init_matches_oracle(gpu, orc):
    from ptychopy_amd._engine import Engine, SCAL_PROBE_MAX
    p = make_problem(m=64, n_side=4, seed=2)
    with Engine(p["n"], p["m"], p["shape"], n_modes=2) as e:
        e.set_frames(p["amp"])
        e.init_probe(beam_size=110e-9, dx_s=2.8159e-9)
        g = e.get_probe()
        pm = e.get_scalars()[SCAL_PROBE_MAX]
        amp = e.get_frames().astype(np.float64)
    ref, refmax = orc.probe_init(amp, 2, 110e-9, 2.8159e-9)
    assert rel_err(g, ref) < 1e-5
    assert pm == pytest.approx(refmax, rel=1e-6)
    assert np.all(g[1] == 0)


def test_positions_are_validated(gpu):
    from ptychopy_amd._engine import Engine, PtykError, PTYK_ERANGE
    with Engine(2, 16, (20, 20)) as e:
        with pytest.raises(PtykError) as ex:
            e.set_positions(np.array([[0, 0], [4.6, 0]], np.float32))  # roundf(4.6)=5 -> 5+16 > 20
        assert ex.value.code == PTYK_ERANGE
        with pytest.raises(PtykError):
            e.set_positions(np.array([[0, 0], [-0.2, 0]], np.float32))
        e.set_positions(np.array([[0, 0], [4.4, 3.5]], np.float32))
        assert e.get_roi().tolist() == [[0, 0], [4, 4]]  # roundf: half away from zero
