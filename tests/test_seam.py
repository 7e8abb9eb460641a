"""The C++ method seam keeps the reference's IPhasingMethod signatures (IPhasingMethod.h:53-69).

A method class written against those signatures (tests/seam/user_method.cpp, with `override` on
every pure virtual) must compile against the driver header; a changed parameter list fails the
compile.  The driver's own ePIE / DM / MLs classes run through the same seam in the GPU driver
tests (tests/test_gpu_driver.py).
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_user_method_compiles_against_reference_seam(tmp_path):
    src = ROOT / "tests" / "seam" / "user_method.cpp"
    cmd = ["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", f"-I{ROOT / 'include'}",
           f"-I{ROOT / 'ptychopy_amd' / 'csrc' / 'host'}", str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_changed_signature_fails_to_compile(tmp_path):
    """The check above has teeth: dropping the fourierErrors parameter no longer overrides."""
    src = (ROOT / "tests" / "seam" / "user_method.cpp").read_text()
    bad = src.replace("std::vector<std::vector<real_t> >& fourierErrors, const std::vector<float2>& scanPositions,",
                      "const std::vector<float2>& scanPositions,", 1)
    bad = bad.replace("if (fourierErrors.size() < iter) fourierErrors.resize(iter);", "")
    assert bad != src
    f = tmp_path / "bad.cpp"
    f.write_text(bad)
    cmd = ["g++", "-std=c++17", "-fsyntax-only", f"-I{ROOT / 'include'}",
           f"-I{ROOT / 'ptychopy_amd' / 'csrc' / 'host'}", str(f)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode != 0 and "override" in r.stderr
