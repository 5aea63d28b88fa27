"""Builds libxfdtd_core_amd.so in-tree (xfdtd_core_amd/lib/).

HIP translation units are compiled by hipcc for gfx950; the host precompute
(plain C++20, no HIP) by g++.  Flags that matter for results:

* ``-ffp-contract=off`` on the FP64 kernels and on the host precompute: no
  fused multiply-adds, so FP64 results equal the reference CPU path bit for
  bit (the reference builds -std=c++20 -O3 without -march).
* ``-Xclang -target-feature -Xclang -packed-fp32-ops`` on every HIP unit:
  keeps the compiler from emitting packed FP32 instructions.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB_DIR = PKG / "lib"
LIB = LIB_DIR / "libxfdtd_core_amd.so"

ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
ARCH = os.environ.get("XFD_OFFLOAD_ARCH", "gfx950")

HIP_COMMON = ["-std=c++20", "-O3", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
              "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument", f"-I{ROOT / 'include'}"]
HOST_COMMON = ["-std=c++20", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
               "-Wno-unused-variable", f"-I{ROOT / 'include'}"]

UNITS = [
    # (source, compiler, extra flags)
    (CSRC / "device" / "kernels_f64.hip", "hip", ["-ffp-contract=off"]),
    (CSRC / "device" / "kernels_f32.hip", "hip", ["-ffp-contract=fast"]),
    (CSRC / "device" / "capi.hip", "hip", ["-ffp-contract=off"]),
    (CSRC / "host" / "simulation.cpp", "host", []),
    (CSRC / "host" / "sim_capi.cpp", "host", []),
    (CSRC / "host" / "nf2ff.cpp", "host", []),
]


def _compile(unit):
    src, kind, extra = unit
    obj = BUILD / (src.stem + ("_hip" if kind == "hip" else "") + ".o")
    if kind == "hip":
        cmd = [HIPCC, *HIP_COMMON, *extra, "-c", str(src), "-o", str(obj)]
    else:
        cmd = ["g++", *HOST_COMMON, *extra, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    """Compiles every unit and links the shared library; returns its path."""
    BUILD.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(UNITS), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, UNITS))
    tmp = LIB.with_suffix(".so.tmp")
    # -rtlib=libgcc: the host code's compiler helpers (complex multiply and
    # divide, __muldc3 / __divdc3) come from libgcc as in a g++ build of the
    # reference, not from clang's runtime, whose complex division rounds
    # differently in the last bits.
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp), "-lpthread",
           "-rtlib=libgcc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


CPP_TESTS = ROOT / "tests" / "cpp"


def build_cpp_programs(verbose: bool = False):
    """Compiles the C++ scenario programs under tests/cpp/ (user code of the
    C++ API, include/xfdtd_core_amd.hpp) with g++ against the library, into
    tests/cpp/build/.  The run path points at the library relative to the
    program, so the tree can move (it does on the way to a GPU host)."""
    out = CPP_TESTS / "build"
    out.mkdir(exist_ok=True)
    built = []
    for src in sorted(CPP_TESTS.glob("*.cpp")):
        exe = out / src.stem
        cmd = ["g++", "-std=c++20", "-O2", "-Wall", f"-I{ROOT / 'include'}", str(src), "-o", str(exe),
               f"-L{LIB_DIR}", "-lxfdtd_core_amd", "-Wl,-rpath,$ORIGIN/../../../xfdtd_core_amd/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        built.append(exe)
        if verbose:
            print(f"built {exe}")
    return built


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
    build_cpp_programs(verbose=True)
    sys.exit(0)
