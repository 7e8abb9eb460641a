"""The engine library loads and exports every symbol the C ABI header declares."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "ptyk.h").read_text()
    return sorted(set(re.findall(r"\b(ptyk_[a-z0-9_]+)\s*\(", text)))


def test_header_lists_the_binding_exports():
    from ptychopy_amd import _engine
    assert set(declared_symbols()) == set(_engine.EXPORTS)


def test_library_exports_every_declared_symbol():
    from ptychopy_amd import _engine
    if not _engine.LIB_PATH.exists():
        pytest.fail("libptyk.so not built (python -m ptychopy_amd.build)")
    lib = ctypes.CDLL(str(_engine.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_engine_refuses_without_library(tmp_path, monkeypatch):
    from ptychopy_amd import _engine
    monkeypatch.setattr(_engine, "LIB_PATH", tmp_path / "nope.so")
    monkeypatch.setattr(_engine, "_lib", None)
    with pytest.raises(RuntimeError, match="missing"):
        _engine.load()
