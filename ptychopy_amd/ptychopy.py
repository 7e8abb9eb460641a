"""The `ptychopy` Python surface on the gfx950 engine.

Same function names, keyword arguments, defaults and return values as the reference
extension module (ptychopy.c:117-140, keyword lists and defaults at ptychopy.c:199-246 and
284-335), driving the C++ IPhaser through include/ptycho.h.  Use as

    from ptychopy_amd import ptychopy
    ptychopy.epienp(diffractionNP=frames, positionNP=pos, size=256, iter=50, ...)

Differences kept on purpose: every call starts from fresh parameters (the reference keeps a
process-wide parameter singleton, so settings leak between calls); the numpy object guess is
copied at its own size (the reference sizes the buffer scanDimsx*scanDimsy, ptychopy.c:402);
``delta_p`` is honoured; random streams are seeded (keyword ``seed``, default 0).
"""
from __future__ import annotations

import ctypes as C
from typing import Any

import numpy as np

from . import _engine

__all__ = ["helloworld", "epie", "dm", "mls", "epienp", "dmnp", "mlsnp", "epiecmdstr", "dmcmdstr", "mlscmdstr",
           "epieinit", "epiepost", "epiestep", "epieresobj", "epieresprobe", "epienpinit", "mlsinit", "mlspost",
           "mlsstep", "mlsresobj", "mlsresprobe"]

# (name, type, default) in the order of the reference keyword list (ptychopy.c:242-246)
KWARGS: tuple[tuple[str, type, Any], ...] = (
    ("jobID", str, ""), ("fp", str, ""), ("fs", int, 0), ("hdf5path", str, ""), ("dpf", int, 1),
    ("beamSize", float, 400e-9), ("probeGuess", str, ""), ("objectGuess", str, ""), ("size", int, 512),
    ("qx", int, 128), ("qy", int, 128), ("nx", int, 256), ("ny", int, 256), ("scanDimsx", int, 26),
    ("scanDimsy", int, 26), ("spiralScan", int, 0), ("flipScanAxis", int, 0), ("mirror1stScanAxis", int, 0),
    ("mirror2ndScanAxis", int, 0), ("stepx", float, 40e-9), ("stepy", float, 40e-9), ("probeModes", int, 1),
    ("lambd", float, 2.3843e-10), ("dx_d", float, 172e-6), ("z", float, 2.2), ("iter", int, 100), ("T", int, 0),
    ("jitterRadius", int, 0), ("delta_p", float, 0.1), ("threshold", int, 0), ("rotate90", int, 0),
    ("sqrtData", int, 0), ("fftShiftData", int, 0), ("binaryOutput", int, 0), ("simulate", int, 0),
    ("phaseConstraint", int, 1), ("updateProbe", int, 10), ("updateModes", int, 20), ("beamstopMask", int, 0),
    ("lf", str, ""), ("PPS", int, 20),
)
NP_ARRAYS = ("diffractionNP", "positionNP", "objectNP", "probeNP")
EXTRA = (("seed", int, 0), ("RMS", int, 0), ("MLC", int, 0))

_bound = False
_phaser = None  # step-mode state (ptychopy.c:113)
_istep = 0


def _lib():
    global _bound
    L = _engine.load()
    if not _bound:
        P, s, i32, i64, u32, d = C.c_void_p, C.c_char_p, C.c_int32, C.c_int64, C.c_uint32, C.c_double
        L.ptycho_phaser_create.argtypes = [C.POINTER(C.c_void_p)]
        L.ptycho_phaser_destroy.argtypes = [P]
        L.ptycho_phaser_destroy.restype = None
        L.ptycho_phaser_last_error.argtypes = [P]
        L.ptycho_phaser_last_error.restype = s
        L.ptycho_phaser_set_param.argtypes = [P, s, s]
        L.ptycho_phaser_parse_cmdline.argtypes = [P, s]
        L.ptycho_phaser_get_param.argtypes = [P, s, C.c_char_p, i32]
        L.ptycho_phaser_set_diffraction.argtypes = [P, P, i64, i32, i32]
        L.ptycho_phaser_set_positions.argtypes = [P, P, i64]
        L.ptycho_phaser_get_fourier_errors.argtypes = [P, C.c_uint32, P, i64]
        L.ptycho_phaser_set_object_guess.argtypes = [P, P, i32, i32]
        L.ptycho_phaser_set_probe_guess.argtypes = [P, P, i32, i32]
        for n in ("init", "phase", "phaseinit", "phasepost"):
            getattr(L, "ptycho_phaser_" + n).argtypes = [P]
        L.ptycho_phaser_add_method.argtypes = [P, s, u32]
        L.ptycho_phaser_phasestep.argtypes = [P, u32]
        L.ptycho_phaser_write_results.argtypes = [P, i32]
        L.ptycho_phaser_object_shape.argtypes = [P, C.POINTER(i32), C.POINTER(i32)]
        L.ptycho_phaser_probe_shape.argtypes = [P, C.POINTER(i32), C.POINTER(i32)]
        L.ptycho_phaser_get_object.argtypes = [P, P]
        L.ptycho_phaser_get_probe.argtypes = [P, P]
        L.ptycho_phaser_num_positions.argtypes = [P]
        L.ptycho_phaser_num_positions.restype = i64
        L.ptycho_phaser_get_positions.argtypes = [P, P, i64]
        L.ptycho_phaser_get_errors.argtypes = [P, P, i32]
        L.ptycho_phaser_phase_seconds.argtypes = [P]
        L.ptycho_phaser_phase_seconds.restype = d
        L.ptycho_phaser_engine.argtypes = [P]
        L.ptycho_phaser_engine.restype = P
        _bound = True
    return L


class Phaser:
    """One C++ IPhaser (include/ptycho.h)."""

    def __init__(self):
        self.L = _lib()
        h = C.c_void_p()
        if self.L.ptycho_phaser_create(C.byref(h)) != 0:
            raise RuntimeError("ptycho_phaser_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.ptycho_phaser_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what}: {self.L.ptycho_phaser_last_error(self.h).decode()}")

    def set_param(self, key: str, value):
        self._chk(self.L.ptycho_phaser_set_param(self.h, key.encode(), str(value).encode()), f"parameter {key}")

    def get_param(self, key: str) -> str:
        buf = C.create_string_buffer(256)
        self._chk(self.L.ptycho_phaser_get_param(self.h, key.encode(), buf, 256), f"parameter {key}")
        return buf.value.decode()

    def parse_cmdline(self, cmd: str):
        self._chk(self.L.ptycho_phaser_parse_cmdline(self.h, cmd.encode()), "command line")

    def set_diffraction(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.ndim != 3:
            raise ValueError("Expected 3 dimensional diffraction array")
        self._chk(self.L.ptycho_phaser_set_diffraction(self.h, a.ctypes.data, a.shape[0], a.shape[1], a.shape[2]),
                  "diffraction")

    def set_positions(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        if p.ndim != 2 or p.shape[1] != 2:
            raise ValueError("Expected N x 2 position array")
        self._chk(self.L.ptycho_phaser_set_positions(self.h, p.ctypes.data, p.shape[0]), "positions")

    def set_object_guess(self, o):
        o = np.ascontiguousarray(o, dtype=np.complex64)
        self._chk(self.L.ptycho_phaser_set_object_guess(self.h, o.ctypes.data, o.shape[0], o.shape[1]), "object")

    def set_probe_guess(self, p):
        p = np.ascontiguousarray(p, dtype=np.complex64)
        self._chk(self.L.ptycho_phaser_set_probe_guess(self.h, p.ctypes.data, p.shape[0], p.shape[1]), "probe")

    def init(self):
        self._chk(self.L.ptycho_phaser_init(self.h), "init")

    def add_method(self, name: str, iterations: int):
        self._chk(self.L.ptycho_phaser_add_method(self.h, name.encode(), int(iterations)), "method")

    def phase(self):
        self._chk(self.L.ptycho_phaser_phase(self.h), "phase")

    def phaseinit(self):
        self._chk(self.L.ptycho_phaser_phaseinit(self.h), "phaseinit")

    def phasestep(self, i: int):
        self._chk(self.L.ptycho_phaser_phasestep(self.h, int(i)), "step")

    def phasepost(self):
        self._chk(self.L.ptycho_phaser_phasepost(self.h), "phasepost")

    def write_results(self, r: int = 0) -> bool:
        return self.L.ptycho_phaser_write_results(self.h, int(r)) == 0

    def get_object(self) -> np.ndarray:
        r, c = C.c_int32(), C.c_int32()
        self._chk(self.L.ptycho_phaser_object_shape(self.h, C.byref(r), C.byref(c)), "object shape")
        out = np.empty((r.value, c.value), np.complex64)
        self._chk(self.L.ptycho_phaser_get_object(self.h, out.ctypes.data), "object")
        return out

    def get_probe(self) -> np.ndarray:
        k, m = C.c_int32(), C.c_int32()
        self._chk(self.L.ptycho_phaser_probe_shape(self.h, C.byref(k), C.byref(m)), "probe shape")
        out = np.empty((k.value * m.value, m.value), np.complex64)
        self._chk(self.L.ptycho_phaser_get_probe(self.h, out.ctypes.data), "probe")
        return out

    def get_positions(self) -> np.ndarray:
        n = self.L.ptycho_phaser_num_positions(self.h)
        out = np.empty((max(n, 0), 2), np.float32)
        self._chk(self.L.ptycho_phaser_get_positions(self.h, out.ctypes.data, n), "positions")
        return out

    def get_errors(self, n: int) -> np.ndarray:
        out = np.zeros(n, np.float64)
        self._chk(self.L.ptycho_phaser_get_errors(self.h, out.ctypes.data, n), "errors")
        return out

    def get_fourier_errors(self, it: int) -> np.ndarray:
        """MLs per-frame Fourier errors of iteration it (filled for it < 10)."""
        n = self.L.ptycho_phaser_num_positions(self.h)
        out = np.zeros(max(n, 0), np.float32)
        self._chk(self.L.ptycho_phaser_get_fourier_errors(self.h, int(it), out.ctypes.data, n), "fourier errors")
        return out

    @property
    def engine_handle(self):
        return self.L.ptycho_phaser_engine(self.h)


def _collect(args, kw, with_arrays: bool):
    names = [k for k, _, _ in KWARGS]
    order = (["jobID"] + list(NP_ARRAYS) + names[1:]) if with_arrays else names
    if len(args) > len(order):
        raise TypeError(f"function takes at most {len(order)} arguments ({len(args)} given)")
    given = dict(zip(order, args))
    for k, v in kw.items():
        if k in given:
            raise TypeError(f"argument for function given by name ('{k}') and position")
        given[k] = v
    allowed = set(order) | {k for k, _, _ in EXTRA}
    for k in given:
        if k not in allowed:
            raise TypeError(f"'{k}' is an invalid keyword argument for this function")
    vals = {}
    for k, t, d in KWARGS + EXTRA:
        v = given.get(k, d)
        try:
            vals[k] = t(v) if t is not str else (v if isinstance(v, str) else str(v))
        except (TypeError, ValueError) as e:
            raise TypeError(f"argument {k}: {e}") from None
    arrays = {k: given.get(k) for k in NP_ARRAYS} if with_arrays else {}
    return vals, arrays


def _make_phaser(algorithm: str, vals: dict, arrays: dict) -> Phaser:
    ph = Phaser()
    ph.set_param("algorithm", algorithm)
    for k, _, _ in KWARGS + EXTRA:
        ph.set_param(k, vals[k])
    if arrays.get("diffractionNP") is not None:
        ph.set_diffraction(arrays["diffractionNP"])
    if arrays.get("positionNP") is not None:
        ph.set_positions(arrays["positionNP"])
    for key, setter, what in (("objectNP", ph.set_object_guess, "object"), ("probeNP", ph.set_probe_guess, "probe")):
        a = arrays.get(key)
        if a is not None:
            a = np.asarray(a)
            if a.ndim != 2:
                raise ValueError(f"Expected 2 dimensional {what} array")
            setter(a)
    return ph


def _run(algorithm: str, args, kw, with_arrays: bool):
    vals, arrays = _collect(args, kw, with_arrays)
    if with_arrays and arrays.get("diffractionNP") is None:
        # ptychopy_algorithmnp prints this and returns early; the entry point still returns
        # True (ptychopy.c:366-367, 523-540)
        print("Please input a diffraction pattern numpy array")
        return True
    ph = _make_phaser(algorithm, vals, arrays)
    try:
        ph.init()
        ph.add_method(algorithm, vals["iter"])
        ph.phase()
        ph.write_results(0)
    finally:
        ph.close()
    return True  # every whole-mode entry point returns Py_True (ptychopy.c:499-540)


def helloworld(list1, list2, list3):
    """Prints elements of 1D/2D/3D float arrays (ptychopy.c:157-195)."""
    a, b, c = (np.asarray(x, np.float64) for x in (list1, list2, list3))
    print(f"1D: {a[2]:f}, 2D: {b[3][1]:f}, 3D: {c[1][0][2]:f}.")


def epie(*args, **kw):
    """ePIE from keyword parameters (ptychopy.c:499-505); returns True."""
    return _run("ePIE", args, kw, False)


def dm(*args, **kw):
    """DM from keyword parameters (ptychopy.c:507-513); returns True."""
    return _run("DM", args, kw, False)


def mls(*args, **kw):
    """MLs from keyword parameters (ptychopy.c:515-521); returns True."""
    return _run("MLs", args, kw, False)


def epienp(*args, **kw):
    """ePIE on numpy inputs diffractionNP / positionNP / objectNP / probeNP (ptychopy.c:523-529);
    returns True."""
    return _run("ePIE", args, kw, True)


def dmnp(*args, **kw):
    """DM on numpy inputs (ptychopy.c:530-535); returns True."""
    return _run("DM", args, kw, True)


def mlsnp(*args, **kw):
    """MLs on numpy inputs (ptychopy.c:537-542); returns True."""
    return _run("MLs", args, kw, True)


def _cmdstr(algorithm: str, cmdstr: str) -> np.ndarray:
    ph = Phaser()
    try:
        ph.parse_cmdline(cmdstr)
        # the entry point names the method; the reference also needs "-algorithm=MLs" in the
        # string for the MLs geometry (IPhaser.cpp:156-166), which is implied here
        ph.set_param("algorithm", algorithm)
        ph.init()
        ph.add_method(algorithm, int(ph.get_param("iterations")))
        ph.phase()
        ph.write_results(0)
        return ph.get_object()
    finally:
        ph.close()


def epiecmdstr(cmdstr: str = "") -> np.ndarray:
    """ePIE from a "-key=value ..." command string; returns the object (ptychopy.c:571-613)."""
    return _cmdstr("ePIE", cmdstr)


def dmcmdstr(cmdstr: str = "") -> np.ndarray:
    return _cmdstr("DM", cmdstr)


def mlscmdstr(cmdstr: str = "") -> np.ndarray:
    return _cmdstr("MLs", cmdstr)


def _step_init(algorithm: str, ph: Phaser, iterations: int):
    global _phaser, _istep
    if _phaser is not None:
        _phaser.close()
    _phaser, _istep = ph, 0
    ph.set_param("algorithm", algorithm)
    ph.init()
    ph.add_method(algorithm, iterations)
    ph.phaseinit()
    return True


def epieinit(cmdstr: str = ""):
    """Step mode: parse, init and prePhase (ptychopy.c:902-936)."""
    ph = Phaser()
    ph.parse_cmdline(cmdstr)
    return _step_init("ePIE", ph, int(ph.get_param("iterations")))


def epienpinit(*args, **kw):
    """Step mode on numpy inputs (ptychopy_algorithmnpinit)."""
    vals, arrays = _collect(args, kw, True)
    if arrays.get("diffractionNP") is None:
        print("Please input a diffraction pattern numpy array")
        return None
    return _step_init("ePIE", _make_phaser("ePIE", vals, arrays), vals["iter"])


def mlsinit(cmdstr: str = ""):
    ph = Phaser()
    ph.parse_cmdline(cmdstr)
    return _step_init("MLs", ph, int(ph.get_param("iterations")))


def _require():
    if _phaser is None:
        raise RuntimeError("call epieinit/epienpinit/mlsinit first")
    return _phaser


def _step():
    global _istep
    _istep += 1
    _require().phasestep(_istep)
    return True


def _post():
    global _phaser
    ph = _require()
    ph.phasepost()
    ph.write_results(0)
    ph.close()
    _phaser = None
    return True


def epiestep():
    return _step()


def mlsstep():
    return _step()


def epiepost():
    return _post()


def mlspost():
    return _post()


def epieresobj() -> np.ndarray:
    return _require().get_object()


def mlsresobj() -> np.ndarray:
    return _require().get_object()


def epieresprobe() -> np.ndarray:
    """All probe modes stacked, (modes*size) x size complex64 (ptychopy.c:978-990)."""
    return _require().get_probe()


def mlsresprobe() -> np.ndarray:
    return _require().get_probe()
