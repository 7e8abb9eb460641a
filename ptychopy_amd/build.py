"""Build the engine shared library (HIP kernels + C++ host driver) for gfx950.

``python -m ptychopy_amd.build`` compiles every source under ``csrc/`` with hipcc and
links ``ptychopy_amd/_lib/libptyk.so``.  Objects go to ``ptychopy_amd/_build/``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIBDIR = PKG / "_lib"
LIB = LIBDIR / "libptyk.so"
INCLUDE = PKG.parent / "include"

ARCH = os.environ.get("PTYK_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# Packed-FP32 VALU forms are disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", f"-I{INCLUDE}", f"-I{CSRC}",
          "-Wno-unused-command-line-argument", "-Wno-unused-result", "-Wno-unused-value"]


def sources() -> list[Path]:
    return sorted(list(CSRC.glob("*.hip")) + list((CSRC / "host").glob("*.cpp")))


def _compile(src: Path) -> Path:
    rel = src.relative_to(CSRC)
    obj = BUILD / (str(rel).replace("/", "__") + ".o")
    obj.parent.mkdir(parents=True, exist_ok=True)
    if src.suffix == ".hip":
        cmd = [HIPCC, *COMMON, *DEVICE_FLAGS, "-c", str(src), "-o", str(obj)]
    else:  # host-only C++ (driver) compiled by hipcc's clang without device code
        cmd = [HIPCC, *COMMON, "-x", "c++", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-c", str(src), "-o",
               str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    """Compile all sources and link the shared library; always rebuilds."""
    BUILD.mkdir(parents=True, exist_ok=True)
    LIBDIR.mkdir(parents=True, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, len(srcs)) or 1
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
