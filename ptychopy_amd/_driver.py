"""ctypes binding of the host driver's C ABI (include/ptycho.h), in the same library."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _engine


class MeshParams(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("step_x", C.c_double), ("step_y", C.c_double),
                ("dx_s", C.c_double), ("drift_x", C.c_double), ("drift_y", C.c_double),
                ("jitter_radius", C.c_int32), ("mirror_x", C.c_int32), ("mirror_y", C.c_int32),
                ("flip", C.c_int32), ("probe_size", C.c_int32), ("reserved", C.c_int32), ("seed", C.c_uint64)]


DRIVER_EXPORTS = (
    "ptycho_sample_pixel_size", "ptycho_mesh_cartesian", "ptycho_mesh_list", "ptycho_mesh_spiral",
    "ptycho_save_complex_csv", "ptycho_load_complex_csv", "ptycho_phaser_create", "ptycho_phaser_destroy",
    "ptycho_phaser_last_error", "ptycho_phaser_set_param", "ptycho_phaser_parse_cmdline", "ptycho_phaser_get_param",
    "ptycho_phaser_set_diffraction", "ptycho_phaser_set_positions", "ptycho_phaser_set_object_guess",
    "ptycho_phaser_set_probe_guess", "ptycho_phaser_init", "ptycho_phaser_add_method", "ptycho_phaser_phase",
    "ptycho_phaser_phaseinit", "ptycho_phaser_phasestep", "ptycho_phaser_phasepost", "ptycho_phaser_write_results",
    "ptycho_phaser_object_shape", "ptycho_phaser_get_object", "ptycho_phaser_probe_shape", "ptycho_phaser_get_probe",
    "ptycho_phaser_num_positions", "ptycho_phaser_get_positions", "ptycho_phaser_get_errors",
    "ptycho_phaser_phase_seconds", "ptycho_phaser_engine", "ptycho_mesh_cartesian_raw", "ptycho_mls_canvas",
    "ptycho_mls_partitions", "ptycho_phaser_get_fourier_errors", "ptycho_mls_probe_modes",
    "ptycho_mls_compact_groups",
)

_bound = False


def lib():
    global _bound
    L = _engine.load()
    if not _bound:
        L.ptycho_sample_pixel_size.argtypes = [C.c_double, C.c_double, C.c_int32, C.c_double]
        L.ptycho_sample_pixel_size.restype = C.c_double
        L.ptycho_mesh_cartesian.argtypes = [C.POINTER(MeshParams), C.c_void_p, C.c_void_p, C.c_void_p]
        L.ptycho_mesh_list.argtypes = [C.POINTER(MeshParams), C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                       C.c_void_p]
        L.ptycho_mesh_spiral.argtypes = [C.POINTER(MeshParams), C.c_void_p, C.c_void_p, C.c_void_p]
        L.ptycho_save_complex_csv.argtypes = [C.c_char_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32]
        L.ptycho_load_complex_csv.argtypes = [C.c_char_p, C.c_void_p, C.c_int32, C.POINTER(C.c_int32),
                                              C.POINTER(C.c_int32)]
        L.ptycho_mesh_cartesian_raw.argtypes = [C.POINTER(MeshParams), C.c_void_p]
        L.ptycho_mls_canvas.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        L.ptycho_mls_partitions.argtypes = [C.c_int64, C.c_int32, C.c_uint64, C.c_void_p, C.c_void_p,
                                            C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.ptycho_mls_compact_groups.argtypes = [C.c_void_p, C.c_int64, C.c_uint64, C.c_void_p, C.c_void_p,
                                                C.POINTER(C.c_int32)]
        L.ptycho_mls_probe_modes.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                             C.c_void_p, C.c_void_p]
        _bound = True
    return L


def mls_compact_groups(pos, seed: int = 0):
    """Compact (k-means) MLs frame groups (IPtychoScanMesh::get_close_indices): list of frame arrays."""
    p = np.ascontiguousarray(pos, dtype=np.float32)
    n = p.shape[0]
    g = C.c_int32()
    L = lib()
    if L.ptycho_mls_compact_groups(p.ctypes.data_as(C.c_void_p), n, seed, None, None, C.byref(g)) != 0:
        raise ValueError("bad positions")
    fr = np.zeros(n, np.int32)
    of = np.zeros(g.value + 1, np.int32)
    if L.ptycho_mls_compact_groups(p.ctypes.data_as(C.c_void_p), n, seed, fr.ctypes.data_as(C.c_void_p),
                                   of.ctypes.data_as(C.c_void_p), C.byref(g)) != 0:
        raise ValueError("bad positions")
    return [fr[of[i]:of[i + 1]] for i in range(g.value)]


def mls_probe_modes(engine, k: int, lambd: float, dx_s: float, beam_size: float, draws, mode0=None) -> np.ndarray:
    """Initial MLs probe modes (Probe::initMLH) with the given draws (2 per extra mode); mode0:
    optional guess for mode 0.  engine: an _engine.Engine of the frame size (its FFT is used)."""
    m = engine.m
    out = np.zeros((k, m, m), np.complex64)
    if mode0 is not None:
        out[0] = mode0
    d = np.ascontiguousarray(draws, dtype=np.float32)
    if lib().ptycho_mls_probe_modes(engine.h, m, k, lambd, dx_s, beam_size, d.ctypes.data_as(C.c_void_p),
                                    out.ctypes.data_as(C.c_void_p)) != 0:
        raise RuntimeError("ptycho_mls_probe_modes failed")
    return out


def sample_pixel_size(lambd: float, z: float, size: int, dx_d: float) -> float:
    """dx_s = lambda z / (size dx_d) rounded to float (Parameters.h:175-179)."""
    return lib().ptycho_sample_pixel_size(lambd, z, size, dx_d)


def cartesian_mesh(nx: int, ny: int, step_x: float, step_y: float, dx_s: float, probe_size: int, *,
                   jitter_radius: int = 0, seed: int = 0, mirror_x=False, mirror_y=False, flip=False,
                   drift=(0.0, 0.0)):
    """Positions (nx*ny, 2) float32 shifted to the canvas, canvas (rows, cols), extents."""
    p = MeshParams(nx, ny, step_x, step_y, dx_s, drift[0], drift[1], jitter_radius, int(mirror_x), int(mirror_y),
                   int(flip), probe_size, 0, seed)
    pos = np.empty((nx * ny, 2), np.float32)
    canvas = np.empty(2, np.int32)
    ext = np.empty(4, np.float32)
    rc = lib().ptycho_mesh_cartesian(C.byref(p), pos.ctypes.data, canvas.ctypes.data, ext.ctypes.data)
    if rc != 0:
        raise ValueError("bad mesh parameters")
    return pos, (int(canvas[0]), int(canvas[1])), ext


def _params(nx, ny, step_x, step_y, dx_s, probe_size, jitter_radius=0, seed=0, mirror_x=False, mirror_y=False,
            flip=False, drift=(0.0, 0.0)):
    return MeshParams(nx, ny, step_x, step_y, dx_s, drift[0], drift[1], jitter_radius, int(mirror_x), int(mirror_y),
                      int(flip), probe_size, 0, seed)


def list_mesh(metres, dx_s: float, probe_size: int, *, jitter_radius=0, seed=0, flip=False, drift=(0.0, 0.0)):
    m = np.ascontiguousarray(metres, dtype=np.float64)
    p = _params(m.shape[0], 1, 0.0, 0.0, dx_s, probe_size, jitter_radius, seed, flip=flip, drift=drift)
    pos = np.empty((m.shape[0], 2), np.float32)
    canvas = np.empty(2, np.int32)
    ext = np.empty(4, np.float32)
    if lib().ptycho_mesh_list(C.byref(p), m.ctypes.data, m.shape[0], pos.ctypes.data, canvas.ctypes.data,
                              ext.ctypes.data) != 0:
        raise ValueError("bad list mesh parameters")
    return pos, (int(canvas[0]), int(canvas[1])), ext


def spiral_mesh(nx: int, ny: int, step_x: float, step_y: float, dx_s: float, probe_size: int, *, jitter_radius=0,
                seed=0, flip=False, drift=(0.0, 0.0)):
    p = _params(nx, ny, step_x, step_y, dx_s, probe_size, jitter_radius, seed, flip=flip, drift=drift)
    pos = np.empty((nx * ny, 2), np.float32)
    canvas = np.empty(2, np.int32)
    ext = np.empty(4, np.float32)
    if lib().ptycho_mesh_spiral(C.byref(p), pos.ctypes.data, canvas.ctypes.data, ext.ctypes.data) != 0:
        raise ValueError("bad spiral mesh parameters")
    return pos, (int(canvas[0]), int(canvas[1])), ext


def save_complex(path: str, a, binary: bool = False) -> bool:
    a = np.ascontiguousarray(a, dtype=np.complex64)
    return lib().ptycho_save_complex_csv(str(path).encode(), a.ctypes.data, a.shape[0], a.shape[1], int(binary)) == 0


def load_complex_csv(path: str) -> np.ndarray:
    r, c = C.c_int32(), C.c_int32()
    if lib().ptycho_load_complex_csv(str(path).encode(), None, 0, C.byref(r), C.byref(c)) != 0:
        raise IOError(path)
    out = np.empty((r.value, c.value), np.complex64)
    if lib().ptycho_load_complex_csv(str(path).encode(), out.ctypes.data, out.size, C.byref(r), C.byref(c)) != 0:
        raise IOError(path)
    return out


def cartesian_mesh_raw(nx: int, ny: int, step_x: float, step_y: float, dx_s: float, probe_size: int, **kw):
    """Unshifted Cartesian positions (nx*ny, 2) float32 in pixels."""
    p = _params(nx, ny, step_x, step_y, dx_s, probe_size, **kw)
    out = np.empty((nx * ny, 2), np.float32)
    if lib().ptycho_mesh_cartesian_raw(C.byref(p), out.ctypes.data) != 0:
        raise ValueError("bad mesh parameters")
    return out


def mls_canvas(pos, extent, probe_size: int):
    """(canvas (rows, cols), box (row0, row1, col0, col1)) of the MLs object for unshifted positions."""
    ps = np.ascontiguousarray(pos, dtype=np.float32)
    ex = np.ascontiguousarray(extent, dtype=np.float32)
    canvas = np.empty(2, np.int32)
    box = np.empty(4, np.int32)
    if lib().ptycho_mls_canvas(ps.ctypes.data, ps.shape[0], ex.ctypes.data, probe_size, canvas.ctypes.data,
                               box.ctypes.data) != 0:
        raise ValueError("bad MLs canvas arguments")
    return (int(canvas[0]), int(canvas[1])), tuple(int(b) for b in box)


def mls_partitions(n: int, niter: int = 100, seed: int = 0):
    """List of partitions; each a list of int32 frame-index arrays (largest group first)."""
    g, p = C.c_int32(), C.c_int32()
    if lib().ptycho_mls_partitions(n, niter, seed, None, None, C.byref(g), C.byref(p)) != 0:
        raise ValueError("bad partition arguments")
    fr = np.empty((p.value, n), np.int32)
    of = np.empty((p.value, g.value + 1), np.int32)
    lib().ptycho_mls_partitions(n, niter, seed, fr.ctypes.data, of.ctypes.data, C.byref(g), C.byref(p))
    return [[fr[i, of[i, j]:of[i, j + 1]] for j in range(g.value)] for i in range(p.value)]
