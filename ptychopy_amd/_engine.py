"""ctypes binding of the engine's C ABI (include/ptyk.h).

The engine is the HIP library ``ptychopy_amd/_lib/libptyk.so``.  There is no CPU fallback:
if the library is missing or no HIP device is present, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libptyk.so"

PTYK_OK = 0
PTYK_EINVAL = -1
PTYK_ENOMEM = -2
PTYK_EDEVICE = -3
PTYK_ESTATE = -4
PTYK_ERANGE = -5

FLAG_PHASE_CONSTRAINT = 1
FLAG_UPDATE_PROBE = 2
FLAG_UPDATE_MODES = 4
FLAG_RMS = 8

SCAL_OBJ_MAX, SCAL_PROBE_MAX, SCAL_ERROR = 0, 1, 2
SCAL_DM_FACTOR_OBJ, SCAL_DM_FACTOR_PROBE, SCAL_COUNT = 4, 5, 8

# Every symbol include/ptyk.h declares.
EXPORTS = (
    "ptyk_create", "ptyk_destroy", "ptyk_last_error", "ptyk_set_frames_f32", "ptyk_set_frames_f16",
    "ptyk_get_frames_f32", "ptyk_get_frame_scales", "ptyk_set_positions", "ptyk_get_roi", "ptyk_set_object",
    "ptyk_get_object", "ptyk_clear_object", "ptyk_set_probe", "ptyk_get_probe", "ptyk_set_psi", "ptyk_get_psi",
    "ptyk_get_scalars", "ptyk_set_scalars", "ptyk_refresh_maxima", "ptyk_init_probe", "ptyk_begin_modal",
    "ptyk_end_modal", "ptyk_epie_iteration", "ptyk_dm_iteration", "ptyk_dm_iteration_timed", "ptyk_fft2",
    "ptyk_simulate", "ptyk_synchronize", "ptyk_set_tuning", "ptyk_get_tuning",
)

TUNE_DM_WAVE = 1
DM_KERNELS = ("row_pass", "column_pass", "inverse_row_pass", "probe_gather", "object_gather", "probe_apply",
              "object_apply", "finish")


class PtykError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ptyk error {code}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [("n_frames", C.c_int32), ("frame_size", C.c_int32), ("n_modes", C.c_int32),
                ("obj_rows", C.c_int32), ("obj_cols", C.c_int32), ("device", C.c_int32),
                ("tex_quantize", C.c_int32), ("reserved", C.c_int32)]


_lib = None


def load():
    """Load libptyk.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m ptychopy_amd.build` first")
    lib = C.CDLL(str(LIB_PATH))
    P, i32, u32, i64, f = C.c_void_p, C.c_int32, C.c_uint32, C.c_int64, C.c_float
    lib.ptyk_create.argtypes = [C.POINTER(Config), C.POINTER(C.c_void_p)]
    lib.ptyk_destroy.argtypes = [P]
    lib.ptyk_destroy.restype = None
    lib.ptyk_last_error.argtypes = [P]
    lib.ptyk_last_error.restype = C.c_char_p
    for name in ("ptyk_set_frames_f32", "ptyk_get_frames_f32", "ptyk_get_frame_scales", "ptyk_set_positions",
                 "ptyk_get_roi", "ptyk_set_object", "ptyk_get_object", "ptyk_set_probe", "ptyk_get_probe",
                 "ptyk_set_psi", "ptyk_get_psi", "ptyk_get_scalars", "ptyk_set_scalars"):
        getattr(lib, name).argtypes = [P, P]
    lib.ptyk_set_frames_f16.argtypes = [P, P, P]
    for name in ("ptyk_clear_object", "ptyk_refresh_maxima", "ptyk_begin_modal", "ptyk_end_modal",
                 "ptyk_synchronize"):
        getattr(lib, name).argtypes = [P]
    lib.ptyk_init_probe.argtypes = [P, C.c_double, C.c_double]
    lib.ptyk_epie_iteration.argtypes = [P, P, u32, C.POINTER(C.c_double)]
    lib.ptyk_dm_iteration.argtypes = [P, u32]
    lib.ptyk_dm_iteration_timed.argtypes = [P, u32, P]
    lib.ptyk_fft2.argtypes = [P, P, i64, C.c_int]
    lib.ptyk_simulate.argtypes = [P, P, i32, i32, P, P, i64, i32, P]
    lib.ptyk_set_tuning.argtypes = [P, i32, i32]
    lib.ptyk_get_tuning.argtypes = [P, i32, C.POINTER(C.c_int32)]
    _lib = lib
    return lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Engine:
    """One reconstruction's device state (a ptyk handle)."""

    def __init__(self, n_frames: int, frame_size: int, obj_shape, n_modes: int = 1, device: int = 0,
                 tex_quantize: int = 0):
        self.lib = load()
        self.n, self.m, self.k = int(n_frames), int(frame_size), int(n_modes)
        self.obj_shape = (int(obj_shape[0]), int(obj_shape[1]))
        cfg = Config(self.n, self.m, self.k, self.obj_shape[0], self.obj_shape[1], int(device), int(tex_quantize), 0)
        h = C.c_void_p()
        rc = self.lib.ptyk_create(C.byref(cfg), C.byref(h))
        if rc != PTYK_OK:
            raise PtykError(rc, self.lib.ptyk_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.ptyk_destroy(self.h)
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

    def _chk(self, rc: int):
        if rc != PTYK_OK:
            raise PtykError(rc, self.lib.ptyk_last_error(self.h).decode())

    # ---- frames
    def set_frames(self, amplitudes):
        a = np.ascontiguousarray(amplitudes, dtype=np.float32)
        assert a.shape == (self.n, self.m, self.m), a.shape
        self._chk(self.lib.ptyk_set_frames_f32(self.h, _ptr(a)))

    def set_frames_f16(self, halves, scales):
        h = np.ascontiguousarray(halves, dtype=np.float16)
        s = np.ascontiguousarray(scales, dtype=np.float32)
        assert h.shape == (self.n, self.m, self.m) and s.shape == (self.n,)
        self._chk(self.lib.ptyk_set_frames_f16(self.h, _ptr(h), _ptr(s)))

    def get_frames(self) -> np.ndarray:
        out = np.empty((self.n, self.m, self.m), np.float32)
        self._chk(self.lib.ptyk_get_frames_f32(self.h, _ptr(out)))
        return out

    def get_frame_scales(self) -> np.ndarray:
        out = np.empty(self.n, np.float32)
        self._chk(self.lib.ptyk_get_frame_scales(self.h, _ptr(out)))
        return out

    # ---- positions
    def set_positions(self, pos):
        p = np.ascontiguousarray(pos, dtype=np.float32)
        assert p.shape == (self.n, 2)
        self._chk(self.lib.ptyk_set_positions(self.h, _ptr(p)))

    def get_roi(self) -> np.ndarray:
        out = np.empty((self.n, 2), np.int32)
        self._chk(self.lib.ptyk_get_roi(self.h, _ptr(out)))
        return out

    # ---- arrays
    def set_object(self, obj):
        o = np.ascontiguousarray(obj, dtype=np.complex64)
        assert o.shape == self.obj_shape
        self._chk(self.lib.ptyk_set_object(self.h, _ptr(o)))

    def get_object(self) -> np.ndarray:
        out = np.empty(self.obj_shape, np.complex64)
        self._chk(self.lib.ptyk_get_object(self.h, _ptr(out)))
        return out

    def clear_object(self):
        self._chk(self.lib.ptyk_clear_object(self.h))

    def set_probe(self, probe):
        p = np.ascontiguousarray(probe, dtype=np.complex64).reshape(self.k, self.m, self.m)
        self._chk(self.lib.ptyk_set_probe(self.h, _ptr(p)))

    def get_probe(self) -> np.ndarray:
        out = np.empty((self.k, self.m, self.m), np.complex64)
        self._chk(self.lib.ptyk_get_probe(self.h, _ptr(out)))
        return out

    def set_psi(self, psi):
        p = np.ascontiguousarray(psi, dtype=np.complex64)
        assert p.shape == (self.n, self.m, self.m)
        self._chk(self.lib.ptyk_set_psi(self.h, _ptr(p)))

    def get_psi(self) -> np.ndarray:
        out = np.empty((self.n, self.m, self.m), np.complex64)
        self._chk(self.lib.ptyk_get_psi(self.h, _ptr(out)))
        return out

    def get_scalars(self) -> np.ndarray:
        out = np.empty(SCAL_COUNT, np.float64)
        self._chk(self.lib.ptyk_get_scalars(self.h, _ptr(out)))
        return out

    def set_scalars(self, s):
        a = np.ascontiguousarray(s, dtype=np.float64)
        assert a.shape == (SCAL_COUNT,)
        self._chk(self.lib.ptyk_set_scalars(self.h, _ptr(a)))

    def refresh_maxima(self):
        self._chk(self.lib.ptyk_refresh_maxima(self.h))

    # ---- solver steps
    def init_probe(self, beam_size: float, dx_s: float):
        self._chk(self.lib.ptyk_init_probe(self.h, float(beam_size), float(dx_s)))

    def begin_modal(self):
        self._chk(self.lib.ptyk_begin_modal(self.h))

    def end_modal(self):
        self._chk(self.lib.ptyk_end_modal(self.h))

    def epie_iteration(self, order, flags: int = 0) -> float:
        o = np.ascontiguousarray(order, dtype=np.int32)
        assert o.shape == (self.n,)
        err = C.c_double(0.0)
        self._chk(self.lib.ptyk_epie_iteration(self.h, _ptr(o), int(flags), C.byref(err)))
        return err.value

    def dm_iteration(self, flags: int = 0):
        self._chk(self.lib.ptyk_dm_iteration(self.h, int(flags)))

    def set_tuning(self, key: int, value: int):
        self._chk(self.lib.ptyk_set_tuning(self.h, int(key), int(value)))

    def get_tuning(self, key: int) -> int:
        v = C.c_int32(0)
        self._chk(self.lib.ptyk_get_tuning(self.h, int(key), C.byref(v)))
        return v.value

    def dm_iteration_timed(self, flags: int = 0) -> np.ndarray:
        ms = np.zeros(len(DM_KERNELS), np.float32)
        self._chk(self.lib.ptyk_dm_iteration_timed(self.h, int(flags), _ptr(ms)))
        return ms

    def synchronize(self):
        self._chk(self.lib.ptyk_synchronize(self.h))

    # ---- utilities
    def fft2(self, x, inverse: bool = False) -> np.ndarray:
        a = np.ascontiguousarray(x, dtype=np.complex64).copy()
        assert a.shape[-2:] == (self.m, self.m)
        n = a.size // (self.m * self.m)
        self._chk(self.lib.ptyk_fft2(self.h, _ptr(a), n, int(inverse)))
        return a

    def simulate(self, obj, probe, pos, squared: bool = False) -> np.ndarray:
        o = np.ascontiguousarray(obj, dtype=np.complex64)
        p = np.ascontiguousarray(probe, dtype=np.complex64)
        ps = np.ascontiguousarray(pos, dtype=np.float32)
        assert p.shape == (self.m, self.m) and ps.ndim == 2 and ps.shape[1] == 2
        out = np.empty((ps.shape[0], self.m, self.m), np.float32)
        self._chk(self.lib.ptyk_simulate(self.h, _ptr(o), o.shape[0], o.shape[1], _ptr(p), _ptr(ps), ps.shape[0],
                                         int(squared), _ptr(out)))
        return out
