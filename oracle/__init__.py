"""FP64 CPU oracle -- test infrastructure only.

Loaded by ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` as the checker.  The engine (``ptychopy_amd``) never imports it.

Thin numpy/ctypes wrappers over ``oracle/_build/liboracle.so`` (built by
``make -C oracle``); see ``oracle.cpp`` for the reference citations of each function.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-C", str(HERE), "-s"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        d, i, i64, f = C.c_double, C.c_int, C.c_int64, C.c_float
        P = C.c_void_p
        _lib.orc_num_threads.restype = i
        _lib.orc_fft2.argtypes = [P, i64, i, i]
        _lib.orc_extract_bilinear.argtypes = [P, i, i, f, f, i, i, P]
        _lib.orc_modulus_projection.argtypes = [P, P, i]
        _lib.orc_probe_init.argtypes = [P, i64, i, i, d, d, P]
        _lib.orc_probe_init.restype = d
        _lib.orc_simulate.argtypes = [P, i, i, P, P, i64, i, i, i, P]
        _lib.orc_epie_iteration.argtypes = [P, P, i, i, P, P, i, i, i, P, P, P, P, i64, i, i, i, i, i, P]
        _lib.orc_dm_iteration.argtypes = [P, i, i, P, P, P, P, i64, i, i, i, P, P]
    return _lib


def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def num_threads() -> int:
    return lib().orc_num_threads()


def fft2(x, inverse: bool = False) -> np.ndarray:
    """Batched 2D FFT of (..., m, m); inverse is unnormalised."""
    a = _c128(x).copy()
    m = a.shape[-1]
    n = a.size // (m * m)
    lib().orc_fft2(_ptr(a), n, m, int(inverse))
    return a


def extract_bilinear(obj, pos, m: int, quantize: int = 0) -> np.ndarray:
    o = _c128(obj)
    out = np.zeros((m, m), np.complex128)
    lib().orc_extract_bilinear(_ptr(o), o.shape[0], o.shape[1], float(np.float32(pos[0])),
                               float(np.float32(pos[1])), m, quantize, _ptr(out))
    return out


def modulus_projection(psi, amp) -> np.ndarray:
    p = _c128(psi).copy()
    a = _f64(amp)
    lib().orc_modulus_projection(_ptr(p), _ptr(a), p.shape[-1])
    return p


def probe_init(amp, k_total: int, beam_size: float, dx_s: float):
    a = _f64(amp)
    n, m, _ = a.shape
    probe = np.zeros((k_total, m, m), np.complex128)
    pmax = lib().orc_probe_init(_ptr(a), n, m, k_total, beam_size, dx_s, _ptr(probe))
    return probe, pmax


def simulate(obj, probe, pos, squared: bool = False, quantize: int = 0) -> np.ndarray:
    o = _c128(obj)
    p = _c128(probe)
    ps = np.ascontiguousarray(pos, dtype=np.float32)
    n = ps.shape[0]
    m = p.shape[-1]
    out = np.zeros((n, m, m), np.float64)
    lib().orc_simulate(_ptr(o), o.shape[0], o.shape[1], _ptr(p), _ptr(ps), n, m, int(squared), quantize, _ptr(out))
    return out


class EpieState:
    """Object, probe and maxima carried between ePIE iterations (all FP64)."""

    def __init__(self, obj, probe, obj_max: float, probe_max: float, obj_int=None, probe_int=None):
        self.obj = _c128(obj).copy()
        self.probe = _c128(probe).copy()
        self.obj_int = _f64(np.abs(self.obj) ** 2 if obj_int is None else obj_int).copy()
        self.probe_int = _f64(np.abs(self.probe) ** 2 if probe_int is None else probe_int).copy()
        self.scal = np.array([obj_max, probe_max, 0.0, 0.0], np.float64)


def epie_iteration(st: EpieState, amp, pos, order, *, k_use: int = 1, modal: bool = False,
                   phase_constraint: bool = False, update_probe: bool = False, rms: bool = False,
                   quantize: int = 0) -> float:
    a = _f64(amp)
    n, m, _ = a.shape
    asum = _f64(a.reshape(n, -1).sum(axis=1))
    ps = np.ascontiguousarray(pos, dtype=np.float32)
    od = np.ascontiguousarray(order, dtype=np.int32)
    k_total = st.probe.shape[0]
    lib().orc_epie_iteration(_ptr(st.obj), _ptr(st.obj_int), st.obj.shape[0], st.obj.shape[1], _ptr(st.probe),
                             _ptr(st.probe_int), k_total, k_use, int(modal), _ptr(a), _ptr(asum), _ptr(ps), _ptr(od),
                             n, m, int(phase_constraint), int(update_probe), int(rms), quantize, _ptr(st.scal))
    return st.scal[2] / n if rms else 0.0


def dm_iteration(obj, probe, psi, amp, pos, factors, *, phase_constraint: bool = False,
                 update_probe: bool = False):
    """One DM iteration; returns (obj, probe, psi, factors, sums) as new arrays."""
    o = _c128(obj).copy()
    p = _c128(probe).copy()
    s = _c128(psi).copy()
    a = _f64(amp)
    ps = np.ascontiguousarray(pos, dtype=np.float32)
    fac = _f64(factors).copy()
    sums = np.zeros(2, np.float64)
    n, m, _ = a.shape
    lib().orc_dm_iteration(_ptr(o), o.shape[0], o.shape[1], _ptr(p), _ptr(s), _ptr(a), _ptr(ps), n, m,
                           int(phase_constraint), int(update_probe), _ptr(fac), _ptr(sums))
    return o, p, s, fac, sums
