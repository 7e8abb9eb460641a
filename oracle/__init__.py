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
        _lib.orc_nn_num_params.argtypes = [i]
        _lib.orc_nn_num_params.restype = i64
        _lib.orc_nn_forward_backward.argtypes = [P, i, i, i, P, P, P, i, P, P, P]
        _lib.orc_nn_conv2d.argtypes = [P, P, P, i, i, i, i, i, i, P]
        _lib.orc_object_error.argtypes = [P, P, P, i, i, P, i64, i, i, d, P]
        _lib.orc_mls_roi.argtypes = [P, i, i, i, i, P, P]
        _lib.orc_mls_init.argtypes = [i, i, i, i, i, P, P, P, P, i, P, P, P, P]
        _lib.orc_mls_ortho_modes.argtypes = [P, i, i]
        _lib.orc_mls_shift.argtypes = [P, i, f, f]
        _lib.orc_mls_lsq.argtypes = [P, P, P, P, P, i, i, P]
        _lib.orc_mls_position_step.argtypes = [P, P, P, i, P, P, P]
        _lib.orc_mls_iteration.argtypes = [C.POINTER(MlsParams), i, i, i, i, i, P, P, P, P, P, d, P, P, P, P, P, P,
                                           P, P, P, i, P]
        _lib.orc_mls_object_norm.argtypes = [P, i, i, P, P, i, i, P, P]
        _lib.orc_mls_remove_extra.argtypes = [i, i, P, P, P]
        _lib.orc_mls_frame_modulus.argtypes = [P, P, i, i]
        _lib.orc_mls_frame_modulus.restype = d
        _lib.orc_mls_variable_probe.argtypes = [i, i, i, P, P, P, i, i, P, P, P, P, P, P]
        _lib.orc_mls_mode_directions.argtypes = [i, i, i, i, i, P, P, P, P, i, d, P, P, P, P, i, i, P, P, P]
        _lib.orc_modulus_modes.argtypes = [P, P, i, i]
        _lib.orc_orthogonalize_modes.argtypes = [P, i, i]
        _lib.orc_init_probe_modes.argtypes = [P, P, i, i]
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


def modulus_modes(psi, amp) -> np.ndarray:
    """Modulus constraint on K exit waves (K x m x m): single-mode form for K = 1, modal for K > 1."""
    p = _c128(psi).copy()
    k, m = p.shape[0], p.shape[-1]
    lib().orc_modulus_modes(_ptr(p), _ptr(_f64(amp)), k, m)
    return p


def orthogonalize_modes(probe) -> np.ndarray:
    """Probe::orthogonalize (reference coefficient convention)."""
    p = _c128(probe).copy()
    lib().orc_orthogonalize_modes(_ptr(p), p.shape[0], p.shape[-1])
    return p


def init_probe_modes(probe):
    """Probe::initProbeModes: (modes, intensities)."""
    p = _c128(probe).copy()
    pi = _f64(np.abs(p) ** 2).copy()
    lib().orc_init_probe_modes(_ptr(p), _ptr(pi), p.shape[0], p.shape[-1])
    return p, pi


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


def object_error(obj, truth, probe, roi, frac: float = 0.01):
    """(e, gamma, pixels) of the normalised object error over the illuminated region."""
    o, t = _c128(obj), _c128(truth)
    p = _c128(probe)
    p = p.reshape(-1, p.shape[-2], p.shape[-1])
    r = np.ascontiguousarray(roi, dtype=np.int32)
    out = np.zeros(4, np.float64)
    lib().orc_object_error(_ptr(o), _ptr(t), _ptr(p), p.shape[0], p.shape[-1], _ptr(r), r.shape[0], o.shape[0],
                           o.shape[1], float(frac), _ptr(out))
    return float(out[0]), complex(out[1], out[2]), int(out[3])


# ---------------------------------------------------------------------------- MLs
class MlsParams(C.Structure):
    """Mirror of oracle_mls.cpp MlsPrm (reference defaults, Parameters.h:158-165)."""
    _fields_ = [("iter", C.c_int32), ("pps", C.c_int32), ("probe_reconstruct", C.c_int32),
                ("object_reconstruct", C.c_int32), ("variable_probe", C.c_int32), ("variable_intensity", C.c_int32),
                ("variable_probe_modes", C.c_int32), ("apply_subpix_shift", C.c_int32),
                ("apply_multimodal_update", C.c_int32), ("n_objects", C.c_int32), ("n_rec", C.c_int32),
                ("reserved", C.c_int32), ("box", C.c_int32 * 4), ("delta_p", C.c_double), ("beta_lsq", C.c_double),
                ("beta_object", C.c_double), ("beta_probe", C.c_double)]


def mls_params(box, **kw) -> MlsParams:
    d = dict(iter=1, pps=20, probe_reconstruct=1, object_reconstruct=1, variable_probe=1, variable_intensity=1,
             variable_probe_modes=1, apply_subpix_shift=1, apply_multimodal_update=0, n_objects=1, n_rec=1,
             delta_p=0.1, beta_lsq=0.9, beta_object=1.0, beta_probe=1.0)
    d.update(kw)
    p = MlsParams()
    for key, v in d.items():
        setattr(p, key, v)
    for i in range(4):
        p.box[i] = int(box[i])
    return p


def mls_roi(pos, m: int, rows: int, cols: int):
    ps = np.ascontiguousarray(pos, dtype=np.float32)
    n = ps.shape[0]
    roi = np.zeros((n, 2), np.int32)
    sub = np.zeros((n, 2), np.float32)
    lib().orc_mls_roi(_ptr(ps), n, m, rows, cols, _ptr(roi), _ptr(sub))
    return roi, sub


def mls_shift(x, sx: float, sy: float) -> np.ndarray:
    """Probe::shift_probe on one frame: sx along columns, sy along rows."""
    a = _c128(x).copy()
    lib().orc_mls_shift(_ptr(a), a.shape[-1], float(np.float32(sx)), float(np.float32(sy)))
    return a


def mls_position_step(op, vp=None, chi=None):
    """(dx, dy, dX, dY) of the gradient position solver for one frame."""
    o = _c128(op)
    m = o.shape[-1]
    v = None if vp is None else _c128(vp)
    c = None if chi is None else _c128(chi)
    out = np.zeros(2, np.float64)
    dX = np.zeros((m, m), np.complex128)
    dY = np.zeros((m, m), np.complex128)
    lib().orc_mls_position_step(_ptr(o), None if v is None else _ptr(v), None if c is None else _ptr(c), m, _ptr(out),
                                _ptr(dX), _ptr(dY))
    return out[0], out[1], dX, dY


def mls_lsq(oup, vp, pum, op, chi) -> np.ndarray:
    """(G, 2) optimal (object, probe) steps of one group."""
    arrs = [_c128(a) for a in (oup, vp, pum, op, chi)]
    g, m = arrs[0].shape[0], arrs[0].shape[-1]
    x = np.zeros((g, 2), np.complex128)
    lib().orc_mls_lsq(*[_ptr(a) for a in arrs], g, m, _ptr(x))
    return x


def mls_object_norm(box, obj, illum, probe, vmode):
    """Probe::calc_object_norm: (obj, probe, vmode) after the renormalisation."""
    o = _c128(obj).copy()
    p = _c128(probe).copy()
    v = _c128(vmode).copy()
    b = np.ascontiguousarray(box, dtype=np.int32)
    lib().orc_mls_object_norm(_ptr(b), o.shape[0], o.shape[1], _ptr(o), _ptr(_f64(illum)), p.shape[0], p.shape[-1],
                              _ptr(p), _ptr(v))
    return o, p, v


def mls_remove_extra(probe, vmode, evo):
    """Probe::remove_extra_degree: (probe, vmode, evo) after it; evo is 2 x n."""
    p = _c128(probe).copy()
    v = _c128(vmode).copy()
    e = _f64(evo).copy()
    lib().orc_mls_remove_extra(e.shape[1], p.shape[-1], _ptr(p), _ptr(v), _ptr(e))
    return p, v, e


def mls_frame_modulus(spectra, amp):
    """(chi, error) of one frame: spectra K x m x m of the exit waves, amp m x m."""
    p = _c128(spectra).copy()
    err = lib().orc_mls_frame_modulus(_ptr(p), _ptr(_f64(amp)), p.shape[0], p.shape[-1])
    return p, float(err)


def mls_variable_probe(frames, roi, illum, pu, pum, op, chi0, vmode, evo1):
    """Variable-mode update of one group: (vmode, evo1) after it."""
    gi = np.ascontiguousarray(frames, dtype=np.int32)
    r = np.ascontiguousarray(roi, dtype=np.int32)
    il = _f64(illum)
    v = _c128(vmode).copy()
    e1 = _f64(evo1).copy()
    g, m = len(gi), v.shape[-1]
    lib().orc_mls_variable_probe(g, e1.shape[0], m, _ptr(gi), _ptr(r), _ptr(il), il.shape[0], il.shape[1],
                                 _ptr(_c128(pu)), _ptr(_c128(pum)), _ptr(_c128(op)), _ptr(_c128(chi0)), _ptr(v),
                                 _ptr(e1))
    return v, e1


def mls_mode_directions(ll, frames, roi, sub, illum, max_illum, chi, op, vp, probe, subpix=False,
                        probe_rec=True, obj_rec=True):
    """(pu, pum, oup) of mode ll (1-based) for one group (MLS.cpp:251-297)."""
    gi = np.ascontiguousarray(frames, dtype=np.int32)
    r = np.ascontiguousarray(roi, dtype=np.int32)
    sb = np.ascontiguousarray(sub, dtype=np.float32)
    il = _f64(illum)
    p = _c128(probe)
    k, m = p.shape[0], p.shape[-1]
    g = len(gi)
    pu = np.zeros((g, m, m), np.complex128)
    pum = np.zeros((m, m), np.complex128)
    oup = np.zeros((g, m, m), np.complex128)
    lib().orc_mls_mode_directions(int(ll), min(k, int(ll)), g, m, int(subpix), _ptr(gi), _ptr(r), _ptr(sb), _ptr(il),
                                  il.shape[1], float(max_illum), _ptr(_c128(chi)), _ptr(_c128(op)), _ptr(_c128(vp)),
                                  _ptr(p), int(probe_rec), int(obj_rec), _ptr(pu), _ptr(pum), _ptr(oup))
    return pu, pum, oup


def mls_ortho_modes(probe) -> np.ndarray:
    p = _c128(probe).copy()
    lib().orc_mls_ortho_modes(_ptr(p), p.shape[0], p.shape[-1])
    return p


class MlsState:
    """Everything an MLs iteration reads and writes (FP64, positions float32)."""

    def __init__(self, obj, probe, vmode, evo1, pos, amp):
        self.obj = _c128(obj).copy()
        self.probe = _c128(probe).copy()
        self.vmode = _c128(vmode).copy()
        self.pos = np.ascontiguousarray(pos, dtype=np.float32).copy()
        self.pos_o = self.pos.copy()
        n = self.pos.shape[0]
        self.evo = np.zeros((2, n), np.float64)
        self.evo[1] = evo1
        self.amp = _f64(amp)
        self.illum = np.zeros(self.obj.shape, np.float64)
        self.max_illum = 0.0
        self.roi, self.sub = mls_roi(self.pos, self.probe.shape[-1], *self.obj.shape)
        self.beta_obj = np.ones(n, np.complex128)
        self.beta_probe = np.ones(n, np.complex128)
        self.ferr = np.zeros(n, np.float64)

    def init(self, frame_power, variable_probe_modes: int = 1):
        """Probe::initEvo; frame_power = per-frame sums of the intensities."""
        k, m, _ = self.probe.shape
        fp = _f64(frame_power)
        mx = np.zeros(1, np.float64)
        rows, cols = self.obj.shape
        lib().orc_mls_init(self.pos.shape[0], m, k, rows, cols, _ptr(self.obj), _ptr(self.probe), _ptr(self.vmode),
                           _ptr(self.evo), variable_probe_modes, _ptr(self.roi), _ptr(fp), _ptr(self.illum), _ptr(mx))
        self.max_illum = float(mx[0])

    def copy(self) -> "MlsState":
        c = object.__new__(MlsState)
        for key, v in self.__dict__.items():
            setattr(c, key, v.copy() if isinstance(v, np.ndarray) else v)
        return c

    def iteration(self, prm: MlsParams, groups):
        """groups: list of frame-index arrays in processing order."""
        gf = np.ascontiguousarray(np.concatenate([np.asarray(g, np.int32) for g in groups]), dtype=np.int32)
        go = np.zeros(len(groups) + 1, np.int32)
        go[1:] = np.cumsum([len(g) for g in groups])
        k, m, _ = self.probe.shape
        rows, cols = self.obj.shape
        lib().orc_mls_iteration(C.byref(prm), self.pos.shape[0], m, k, rows, cols, _ptr(self.obj), _ptr(self.probe),
                                _ptr(self.vmode), _ptr(self.evo), _ptr(self.illum), self.max_illum, _ptr(self.pos),
                                _ptr(self.pos_o), _ptr(self.roi), _ptr(self.sub), _ptr(self.beta_obj),
                                _ptr(self.beta_probe), _ptr(self.amp), _ptr(gf), _ptr(go), len(groups),
                                _ptr(self.ferr))


def preprocess_frames(raw, size: int, center=None, rotate90: int = 0, fft_shift: bool = False,
                      threshold_sqrt: bool = False, square_root: bool = False, threshold: float = 0.0):
    """DiffractionLoader::run (DiffractionLoader.cpp:55-131) with numpy's own routines: the size x size
    window of each raw frame centred at `center` (rows/cols outside the frame are zero; a frame that
    is already size x size is kept whole), np.rot90 once when rotate90 % 4 == 1 (d_rot90,
    utilitiesKernels.cu:841-851, 2031-2049), np.fft.fftshift when fft_shift (h_shiftFFT), then the
    optional d_preprocessDiffractions step in float32 (DiffractionsKernels.cu:109-118)."""
    r = np.asarray(raw, np.float32)
    if r.ndim == 2:
        r = r[None]
    cnt, rows, cols = r.shape
    cr, cc = center if center is not None else (rows // 2, cols // 2)
    out = np.zeros((cnt, size, size), np.float32)
    if rows == size and cols == size:
        out[:] = r
    else:
        r0, c0 = cr - size // 2, cc - size // 2
        a0, a1 = max(r0, 0), min(r0 + size, rows)
        b0, b1 = max(c0, 0), min(c0 + size, cols)
        if a1 > a0 and b1 > b0:
            out[:, a0 - r0:a1 - r0, b0 - c0:b1 - c0] = r[:, a0:a1, b0:b1]
    if rotate90 % 4 == 1:
        out = np.rot90(out, k=1, axes=(1, 2))
    if fft_shift:
        out = np.fft.fftshift(out, axes=(1, 2))
    out = np.ascontiguousarray(out)
    if threshold_sqrt:
        t = np.float32(threshold)
        v = np.where(np.isfinite(out), out, np.float32(0))
        v = np.where(v < t, np.float32(0), (v - t).astype(np.float32))
        out = np.sqrt(v).astype(np.float32) if square_root else v.astype(np.float32)
    return out


def mls_zone_plate_probe(m: int, lambd: float, dx_s: float, beam_size: float) -> np.ndarray:
    """Mode 0 of Probe::initMLH without a guess (h_initProbeMLH / d_initProbeMLH / h_endProbeMLH,
    ProbeKernels.cu:220-366, 592-641), FP64 with numpy's FFT.  Grid x_t = -dx M/2 + t dx M/(M-1)
    (t along columns, the same along rows), aperture D_FZP/2 >= r >= D_H/2 (the host passes the
    diameters halved), kernel exp(i pi fzp r^2) exp(i k r^2 / 2z) placed fftshifted; forward FFT;
    times exp(i k z) exp(i k (u_x^2 + u_y^2) / 2z) with the frequency grid split at M/2; shifted back."""
    rn, drn = 90e-6, 50e-9
    fl = 2 * rn * drn / lambd
    d_fzp, d_h = 180e-6, 60e-6
    dx_fzp = lambd * fl / (m * dx_s)
    fzp = -2 / (lambd * 2 * fl) * np.pi
    idx = np.arange(m)
    x = -dx_fzp * m / 2 + dx_fzp * m / (m - 1) * idx
    xx, yy = np.meshgrid(x, x)  # xx varies along columns, yy along rows
    r2 = xx ** 2 + yy ** 2
    r = np.sqrt(r2)
    k = 2 * np.pi / lambd
    z = fl + beam_size
    inside = (r <= d_fzp / 2) & (r >= d_h / 2)
    kern = np.where(inside, np.exp(1j * (fzp * r2 + k / (2 * z) * r2)), 0)
    fu = lambd * (fl + beam_size) / dx_fzp
    step = fu / (m - 1)
    u = np.where(idx < m // 2, -fu / 2 + step * (m // 2) + step * idx, -fu / 2 + step * (idx - m // 2))
    ux, uy = np.meshgrid(u, u)
    pf = np.exp(1j * k * z) * np.exp(1j * k / (2 * z) * (ux ** 2 + uy ** 2))
    f = np.fft.fft2(np.fft.fftshift(kern))
    return np.fft.ifftshift(f * pf)


def mls_probe_modes(m: int, k: int, lambd: float, dx_s: float, beam_size: float, draws, mode0=None) -> np.ndarray:
    """Probe::initMLH (Probe.cpp:376-494) in FP64: mode 0 (a guess or the zone-plate probe); modes
    i >= 1 = mode 0 exp(-2 pi i (r1 (c/M - 0.5) + r2 (r/M - 0.5))) (h_initVarProbe with
    d_imshift_fft, ProbeKernels.cu:756-772, utilitiesKernels.cu:784-826), (r1, r2) = draws[2i-2:2i];
    then the eigenmodes of A_ij = sum conj(P_i) P_j (h_Orth / d_innerProductOne,
    utilitiesKernels.cu:870-883), out_e = sum_r P_r v_re (d_innerProductModes,
    utilitiesKernels.cu:888-907), strongest |eigenvalue| first, each eigenvector scaled so its
    largest component is real positive (the reference leaves Eigen's order and phases)."""
    p0 = mls_zone_plate_probe(m, lambd, dx_s, beam_size) if mode0 is None else _c128(mode0)
    modes = np.zeros((k, m, m), np.complex128)
    modes[0] = p0
    d = np.asarray(draws, np.float64)
    c = np.arange(m) / m - 0.5
    for i in range(1, k):
        r1, r2 = d[2 * i - 2], d[2 * i - 1]
        modes[i] = p0 * np.exp(-2j * np.pi * (r1 * c[None, :] + r2 * c[:, None]))
    if k == 1:
        return modes
    flat = modes.reshape(k, -1)
    a = np.conj(flat) @ flat.T
    w, v = np.linalg.eigh(a)
    order = sorted(range(k), key=lambda e: -abs(w[e]))
    out = np.zeros_like(flat)
    for o, e in enumerate(order):
        vec = v[:, e]
        big = int(np.argmax(np.abs(vec)))
        vec = vec * np.conj(vec[big]) / abs(vec[big])
        out[o] = vec @ flat
    return out.reshape(k, m, m)


# ---------------------------------------------------------------------------- warm-start network
def nn_num_params(c: int) -> int:
    return int(lib().orc_nn_num_params(c))


def nn_forward_backward(params, x, t, c: int, emulate_bf16: bool = False, grad: bool = True, weight=None):
    """(out, loss, grad) of the encoder-decoder; x: batch x s x s, t: batch x 2 x s x s;
    weight: optional s x s per-pixel loss weight."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    wa = None if weight is None else np.ascontiguousarray(weight, dtype=np.float32)
    xa = np.ascontiguousarray(x, dtype=np.float32)
    ta = np.ascontiguousarray(t, dtype=np.float32)
    b, s = xa.shape[0], xa.shape[-1]
    out = np.zeros((b, 2, s, s), np.float64)
    loss = C.c_double(0.0)
    g = np.zeros(p.size, np.float64) if grad else None
    lib().orc_nn_forward_backward(_ptr(p), s, c, b, _ptr(xa), _ptr(ta), None if wa is None else _ptr(wa),
                                  int(emulate_bf16), _ptr(out), C.byref(loss), None if g is None else _ptr(g))
    return out, loss.value, g


def nn_conv2d(x, w, b, relu: bool = False):
    xa = _f64(x)
    wa = _f64(w)
    ba = _f64(b)
    n, cin, h, _ = xa.shape
    cout, _, ks, _ = wa.shape
    y = np.zeros((n, cout, h, h), np.float64)
    lib().orc_nn_conv2d(_ptr(xa), _ptr(wa), _ptr(ba), n, cin, cout, ks, h, int(relu), _ptr(y))
    return y
