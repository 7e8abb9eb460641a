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
    "ptyk_end_modal", "ptyk_orthogonalize_modes", "ptyk_epie_iteration", "ptyk_dm_iteration", "ptyk_dm_iteration_timed", "ptyk_fft2",
    "ptyk_simulate", "ptyk_simulate_noalloc", "ptyk_preprocess_frames", "ptyk_synchronize", "ptyk_set_tuning", "ptyk_get_tuning", "ptyk_mls_begin",
    "ptyk_mls_iteration", "ptyk_mls_iteration_timed", "ptyk_mls_get_state", "ptyk_mls_set_state", "ptyk_set_true_object", "ptyk_object_error",
    "ptyk_batch_create", "ptyk_batch_destroy", "ptyk_batch_last_error", "ptyk_batch_dm_iteration",
    "ptyk_batch_synchronize", "ptyk_batch_graph_count", "ptyk_net_create", "ptyk_net_destroy",
    "ptyk_net_last_error", "ptyk_net_num_params", "ptyk_net_get_params", "ptyk_net_set_params",
    "ptyk_net_set_loss_weight", "ptyk_net_set_data", "ptyk_net_reserve_data", "ptyk_net_put_data", "ptyk_net_train",
    "ptyk_net_evaluate", "ptyk_net_predict", "ptyk_net_predict_amplitudes", "ptyk_net_gradient",
    "ptyk_stitch_patches",
)

TUNE_DM_WAVE = 1
TUNE_MLS_MOD_CB = 2
TUNE_MLS_PUM_FPC = 3
TUNE_MLS_GRAPHS = 4
TUNE_DM_GRAPH = 5
TUNE_MLS_ROW_WAVES = 6
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


class MlsConfig(C.Structure):
    """ptyk_mls_config (include/ptyk.h); defaults are the reference's (Parameters.h:158-165)."""
    _fields_ = [("probe_pos_search", C.c_int32), ("probe_reconstruct", C.c_int32),
                ("object_reconstruct", C.c_int32), ("variable_probe", C.c_int32),
                ("variable_intensity", C.c_int32), ("variable_probe_modes", C.c_int32),
                ("apply_subpix_shift", C.c_int32), ("apply_multimodal_update", C.c_int32),
                ("n_objects", C.c_int32), ("n_rec", C.c_int32), ("object_box", C.c_int32 * 4),
                ("delta_p", C.c_double), ("beta_lsq", C.c_double), ("beta_object", C.c_double),
                ("beta_probe", C.c_double), ("reserved", C.c_int32 * 4)]


MLS_DEFAULTS = dict(probe_pos_search=20, probe_reconstruct=1, object_reconstruct=1, variable_probe=1,
                    variable_intensity=1, variable_probe_modes=1, apply_subpix_shift=1, apply_multimodal_update=0,
                    n_objects=1, n_rec=1, delta_p=0.1, beta_lsq=0.9, beta_object=1.0, beta_probe=1.0)


def mls_config(box, **kw) -> MlsConfig:
    d = dict(MLS_DEFAULTS)
    for key in kw:
        if key not in d:
            raise KeyError(key)
    d.update(kw)
    c = MlsConfig()
    for key, v in d.items():
        setattr(c, key, v)
    for i in range(4):
        c.object_box[i] = int(box[i])
    return c


class IngestParams(C.Structure):
    _fields_ = [("raw_rows", C.c_int32), ("raw_cols", C.c_int32), ("center_row", C.c_int32),
                ("center_col", C.c_int32), ("rotate90", C.c_int32), ("fft_shift", C.c_int32),
                ("threshold_sqrt", C.c_int32), ("square_root", C.c_int32), ("threshold", C.c_float),
                ("reserved", C.c_int32 * 3)]


class MlsState(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("positions_start", C.c_void_p), ("roi", C.c_void_p),
                ("subpixel", C.c_void_p), ("evolution", C.c_void_p), ("variable_mode", C.c_void_p),
                ("illumination", C.c_void_p), ("beta_object", C.c_void_p), ("beta_probe", C.c_void_p),
                ("fourier_errors", C.c_void_p), ("max_illumination", C.c_double)]


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
                 "ptyk_orthogonalize_modes",
                 "ptyk_synchronize"):
        getattr(lib, name).argtypes = [P]
    lib.ptyk_init_probe.argtypes = [P, C.c_double, C.c_double]
    lib.ptyk_epie_iteration.argtypes = [P, P, u32, C.POINTER(C.c_double)]
    lib.ptyk_dm_iteration.argtypes = [P, u32]
    lib.ptyk_dm_iteration_timed.argtypes = [P, u32, P]
    lib.ptyk_fft2.argtypes = [P, P, i64, C.c_int]
    lib.ptyk_simulate.argtypes = [P, P, i32, i32, P, P, i64, i32, P]
    lib.ptyk_simulate_noalloc.argtypes = [P, P, i32, i32, P, P, i64, i32, P]
    lib.ptyk_preprocess_frames.argtypes = [P, P, i64, C.POINTER(IngestParams), P]
    lib.ptyk_set_tuning.argtypes = [P, i32, i32]
    lib.ptyk_get_tuning.argtypes = [P, i32, C.POINTER(C.c_int32)]
    lib.ptyk_mls_begin.argtypes = [P, C.POINTER(MlsConfig), P, P, P, P]
    lib.ptyk_mls_iteration.argtypes = [P, i32, P, P, i32, P]
    lib.ptyk_mls_iteration_timed.argtypes = [P, i32, P, P, i32, P, P]
    lib.ptyk_mls_get_state.argtypes = [P, C.POINTER(MlsState)]
    lib.ptyk_mls_set_state.argtypes = [P, C.POINTER(MlsState)]
    lib.ptyk_set_true_object.argtypes = [P, P]
    lib.ptyk_batch_create.argtypes = [P, i32, C.POINTER(C.c_void_p)]
    lib.ptyk_batch_destroy.argtypes = [P]
    lib.ptyk_batch_destroy.restype = None
    lib.ptyk_batch_last_error.argtypes = [P]
    lib.ptyk_batch_last_error.restype = C.c_char_p
    lib.ptyk_batch_dm_iteration.argtypes = [P, u32]
    lib.ptyk_batch_synchronize.argtypes = [P]
    lib.ptyk_batch_graph_count.argtypes = [P]
    lib.ptyk_object_error.argtypes = [P, C.c_double, P]
    lib.ptyk_net_create.argtypes = [i32, i32, i32, i32, C.c_uint64, C.POINTER(C.c_void_p)]
    lib.ptyk_net_destroy.argtypes = [P]
    lib.ptyk_net_destroy.restype = None
    lib.ptyk_net_last_error.argtypes = [P]
    lib.ptyk_net_last_error.restype = C.c_char_p
    lib.ptyk_net_num_params.argtypes = [P]
    lib.ptyk_net_num_params.restype = i64
    lib.ptyk_net_get_params.argtypes = [P, P]
    lib.ptyk_net_set_params.argtypes = [P, P]
    lib.ptyk_net_set_data.argtypes = [P, P, P, i64]
    lib.ptyk_net_set_loss_weight.argtypes = [P, P]
    lib.ptyk_net_reserve_data.argtypes = [P, i64]
    lib.ptyk_net_put_data.argtypes = [P, i64, P, P, i64]
    lib.ptyk_net_train.argtypes = [P, P, i64, f, C.POINTER(C.c_double)]
    lib.ptyk_net_evaluate.argtypes = [P, P, i64, C.POINTER(C.c_double)]
    lib.ptyk_net_predict.argtypes = [P, P, i64, P]
    lib.ptyk_net_predict_amplitudes.argtypes = [P, P, i64, P]
    lib.ptyk_stitch_patches.argtypes = [P, P, P, i64, f, f, f]
    lib.ptyk_net_gradient.argtypes = [P, P, P, C.POINTER(C.c_double), P]
    _lib = lib
    return lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _group_tables(groups):
    gf = np.ascontiguousarray(np.concatenate([np.asarray(g, np.int32) for g in groups]), dtype=np.int32)
    go = np.zeros(len(groups) + 1, np.int32)
    go[1:] = np.cumsum([len(g) for g in groups])
    return gf, go


# launch kinds of ptyk_mls_iteration_timed, in order (include/ptyk.h)
MLS_KINDS = ("iteration_steps", "illumination_probes", "exit_row_pass", "modulus_column_pass", "inverse_row_pass",
             "probe_update", "object_direction", "lsq", "step_application", "probe_modes", "side_stream_wait",
             "exit_row_pass_own", "modulus_column_pass_own", "inverse_row_pass_own")


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

    def orthogonalize_modes(self):
        """Probe::orthogonalize on the handle's probe modes."""
        self._chk(self.lib.ptyk_orthogonalize_modes(self.h))

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

    def preprocess_frames(self, raw, center=None, rotate90: int = 0, fft_shift: bool = False,
                          threshold_sqrt: bool = False, square_root: bool = False, threshold: float = 0.0):
        """Raw frames (count x rows x cols) -> preprocessed count x M x M frames (ptyk_preprocess_frames)."""
        r = np.ascontiguousarray(raw, dtype=np.float32)
        if r.ndim == 2:
            r = r[None]
        cnt, rows, cols = r.shape
        cr, cc = center if center is not None else (rows // 2, cols // 2)
        p = IngestParams(rows, cols, int(cr), int(cc), int(rotate90), int(bool(fft_shift)), int(bool(threshold_sqrt)),
                         int(bool(square_root)), float(threshold))
        out = np.empty((cnt, self.m, self.m), np.float32)
        self._chk(self.lib.ptyk_preprocess_frames(self.h, _ptr(r), cnt, C.byref(p), _ptr(out)))
        return out

    # ---- MLs
    def mls_begin(self, cfg: MlsConfig, positions, variable_mode, evolution_draws, frame_power=None):
        p = np.ascontiguousarray(positions, dtype=np.float32)
        v = np.ascontiguousarray(variable_mode, dtype=np.complex64)
        e = np.ascontiguousarray(evolution_draws, dtype=np.float64)
        assert p.shape == (self.n, 2) and v.shape == (self.m, self.m) and e.shape == (self.n,)
        fp = None if frame_power is None else np.ascontiguousarray(frame_power, dtype=np.float64)
        self._chk(self.lib.ptyk_mls_begin(self.h, C.byref(cfg), _ptr(p), _ptr(v), _ptr(e),
                                          None if fp is None else _ptr(fp)))

    def mls_iteration(self, it: int, groups):
        """groups: sequence of frame-index arrays in processing order; returns Fourier errors."""
        gf, go = _group_tables(groups)
        ferr = np.zeros(self.n, np.float32)
        self._chk(self.lib.ptyk_mls_iteration(self.h, int(it), _ptr(gf), _ptr(go), len(groups), _ptr(ferr)))
        return ferr

    def mls_iteration_timed(self, it: int, groups):
        """The same iteration with HIP-event timing of its main-stream launches: a dict of ms per
        launch kind (MLS_KINDS) plus 'host_enqueue' (host time to enqueue the iteration)."""
        gf, go = _group_tables(groups)
        ferr = np.zeros(self.n, np.float32)
        ms = np.zeros(len(MLS_KINDS) + 1, np.float32)
        self._chk(self.lib.ptyk_mls_iteration_timed(self.h, int(it), _ptr(gf), _ptr(go), len(groups), _ptr(ferr),
                                                    _ptr(ms)))
        out = {k: float(v) for k, v in zip(MLS_KINDS, ms)}
        out["host_enqueue"] = float(ms[-1])
        return out

    def mls_get_state(self) -> dict:
        n, m = self.n, self.m
        out = dict(positions=np.empty((n, 2), np.float32), positions_start=np.empty((n, 2), np.float32),
                   roi=np.empty((n, 2), np.int32), subpixel=np.empty((n, 2), np.float32),
                   evolution=np.empty((2, n), np.float64), variable_mode=np.empty((m, m), np.complex64),
                   illumination=np.empty(self.obj_shape, np.float32), beta_object=np.empty(n, np.complex64),
                   beta_probe=np.empty(n, np.complex64), fourier_errors=np.empty(n, np.float32))
        st = MlsState(**{k: _ptr(v) for k, v in out.items()}, max_illumination=0.0)
        self._chk(self.lib.ptyk_mls_get_state(self.h, C.byref(st)))
        out["max_illumination"] = st.max_illumination
        return out

    def mls_set_state(self, **arrays):
        """Copy in any of the mls_get_state arrays (and max_illumination)."""
        kinds = dict(positions=np.float32, positions_start=np.float32, roi=np.int32, subpixel=np.float32,
                     evolution=np.float64, variable_mode=np.complex64, illumination=np.float32,
                     beta_object=np.complex64, beta_probe=np.complex64, fourier_errors=np.float32)
        keep = {}
        st = MlsState()
        for key, v in arrays.items():
            if key == "max_illumination":
                st.max_illumination = float(v)
                continue
            a = np.ascontiguousarray(v, dtype=kinds[key])
            keep[key] = a
            setattr(st, key, _ptr(a))
        self._chk(self.lib.ptyk_mls_set_state(self.h, C.byref(st)))

    # ---- object error
    def set_true_object(self, truth):
        t = np.ascontiguousarray(truth, dtype=np.complex64)
        assert t.shape == self.obj_shape
        self._chk(self.lib.ptyk_set_true_object(self.h, _ptr(t)))

    def stitch_patches(self, patches, frames, rel_threshold: float = 1e-3, fill=1.0 + 0j):
        """Object <- probe-intensity-weighted average of (amplitude, phase) patches at the frames' ROIs."""
        p = np.ascontiguousarray(patches, dtype=np.float32)
        fr = np.ascontiguousarray(frames, dtype=np.int32)
        assert p.shape == (fr.size, 2, self.m, self.m)
        self._chk(self.lib.ptyk_stitch_patches(self.h, _ptr(p), _ptr(fr), fr.size, float(rel_threshold),
                                               float(np.real(fill)), float(np.imag(fill))))

    def object_error(self, frac: float = 0.01):
        """(e, gamma, pixels): min_gamma ||gamma O - T|| / ||T|| over the illuminated region."""
        out = np.zeros(4, np.float64)
        self._chk(self.lib.ptyk_object_error(self.h, float(frac), _ptr(out)))
        return float(out[0]), complex(out[1], out[2]), int(out[3])


class Batch:
    """Several Engines advanced together: one captured HIP graph per DM iteration flag set."""

    def __init__(self, engines):
        self.lib = load()
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e.h.value for e in self.engines])
        h = C.c_void_p()
        rc = self.lib.ptyk_batch_create(arr, len(self.engines), C.byref(h))
        if rc != PTYK_OK:
            raise PtykError(rc, self.lib.ptyk_last_error(None).decode())
        self.h = h

    def _chk(self, rc):
        if rc != PTYK_OK:
            raise PtykError(rc, self.lib.ptyk_batch_last_error(self.h).decode())

    def dm_iteration(self, flags: int = 0):
        self._chk(self.lib.ptyk_batch_dm_iteration(self.h, int(flags)))

    def synchronize(self):
        self._chk(self.lib.ptyk_batch_synchronize(self.h))

    @property
    def graph_count(self) -> int:
        return self.lib.ptyk_batch_graph_count(self.h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.ptyk_batch_destroy(self.h)
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
