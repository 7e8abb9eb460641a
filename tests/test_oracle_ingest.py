"""Frame preprocessing oracle (oracle.preprocess_frames, numpy routines) against the reference's own
index formulas, restated per element from its kernels:

  d_symmetrizeDiffraction with the host offsets of DiffractionLoader::run (DiffractionLoader.cpp:84-106)
  d_rot90: dst[(rows - col - 1) pitch + row] = src[row pitch + col] (utilitiesKernels.cu:841-851)
  d_shiftX / d_shiftY: out[(i - round(n/2)) mod n] = in[i] per axis (utilitiesKernels.cu:737-781)
  d_preprocessDiffractions (DiffractionsKernels.cu:109-118)
"""
import math

import numpy as np
import pytest


def _ref_symmetrize(raw, S, center):
    """The reference's copy (in-bounds cases only: the window must not start before row/col 0 while
    ending inside the frame, where the reference writes past its S x S output)."""
    rows, cols = raw.shape
    out = np.zeros((S, S), np.float32)
    if rows == S and cols == S:
        return raw.copy()
    dsx, dsy = center[0] - S // 2, center[1] - S // 2
    out_x, data_x = (abs(dsx), 0) if dsx < 0 else (0, dsx)
    out_y, data_y = (abs(dsy), 0) if dsy < 0 else (0, dsy)
    ex, ey = center[0] + S // 2, center[1] + S // 2
    size_x = S if ex < rows else rows - data_x
    size_y = S if ey < cols else cols - data_y
    for bx in range(size_x):
        for ty in range(size_y):
            out[out_x + bx, out_y + ty] = raw[data_x + bx, data_y + ty]
    return out


def _ref_rot90(a):
    S = a.shape[0]
    out = np.zeros_like(a)
    for row in range(S):
        for col in range(S):
            out[S - col - 1, row] = a[row, col]
    return out


def _ref_shift(a):
    n0, n1 = a.shape
    t = np.zeros_like(a)
    for i in range(n0):
        t[(i - int(round(n0 / 2))) % n0] = a[i]
    out = np.zeros_like(a)
    for j in range(n1):
        out[:, (j - int(round(n1 / 2))) % n1] = t[:, j]
    return out


@pytest.mark.parametrize("raw_shape,center", [((12, 10), (6, 5)), ((20, 16), (9, 11)), ((6, 8), (3, 4)),
                                              ((8, 8), (2, 6)), ((10, 14), (6, 10))])
@pytest.mark.parametrize("rotate,shift", [(0, 0), (1, 0), (0, 1), (1, 1), (2, 1), (5, 0)])
def test_preprocess_matches_reference_index_formulas(orc, raw_shape, center, rotate, shift):
    S = 8
    rng = np.random.default_rng(sum(raw_shape) + 7 * rotate + shift)
    raw = rng.uniform(0, 100, raw_shape).astype(np.float32)
    expect = _ref_symmetrize(raw, S, center)
    if rotate % 4 == 1:
        expect = _ref_rot90(expect)
    if shift:
        expect = _ref_shift(expect)
    got = orc.preprocess_frames(raw, S, center=center, rotate90=rotate, fft_shift=bool(shift))[0]
    assert np.array_equal(got, expect)


def test_preprocess_square_frame_ignores_the_centre(orc):
    raw = np.arange(64, dtype=np.float32).reshape(8, 8)
    assert np.array_equal(orc.preprocess_frames(raw, 8, center=(1, 7))[0], raw)


def test_preprocess_window_before_the_first_row_is_clipped(orc):
    """Centre 2 rows from the top with the window ending inside a 20-row frame: the reference's copy
    would run 2 rows past its 8 x 8 output; here the window keeps raw rows -2..5 (the first two zero)."""
    raw = np.arange(20 * 8, dtype=np.float32).reshape(20, 8)
    got = orc.preprocess_frames(raw, 8, center=(2, 4))[0]
    assert np.array_equal(got[2:], raw[0:6])
    assert not got[:2].any()


def test_threshold_and_square_root(orc):
    raw = np.array([[-1.0, np.nan, np.inf, 2.0], [5.0, 2.5, 0.0, 102.0]], np.float32)
    out = orc.preprocess_frames(np.pad(raw, ((0, 2), (0, 0))), 4, center=(2, 2), threshold_sqrt=True,
                                square_root=True, threshold=2.0)[0]
    assert np.array_equal(out[0], np.array([0, 0, 0, 0], np.float32))
    assert np.array_equal(out[1], np.sqrt(np.array([3.0, 0.5, 0.0, 100.0], np.float32)))
    plain = orc.preprocess_frames(np.pad(raw, ((0, 2), (0, 0))), 4, center=(2, 2), threshold_sqrt=True,
                                  threshold=2.0)[0]
    assert np.array_equal(plain[1], np.array([3.0, 0.5, 0.0, 100.0], np.float32))
    assert math.isclose(float(out[1, 3]), 10.0)
