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


DRIVER_EXPORTS = ("ptycho_sample_pixel_size", "ptycho_mesh_cartesian")

_bound = False


def lib():
    global _bound
    L = _engine.load()
    if not _bound:
        L.ptycho_sample_pixel_size.argtypes = [C.c_double, C.c_double, C.c_int32, C.c_double]
        L.ptycho_sample_pixel_size.restype = C.c_double
        L.ptycho_mesh_cartesian.argtypes = [C.POINTER(MeshParams), C.c_void_p, C.c_void_p, C.c_void_p]
        _bound = True
    return L


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
