import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def _ensure_built():
    """Builds the oracle and the product library when they are missing.

    Builds happen only where the sources and compilers are (the development
    container); a GPU run uses the libraries that travel with the tree.
    """
    oracle = ROOT / "oracle" / "build" / "librefcpu.so"
    if not oracle.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    lib = ROOT / "xfdtd_core_amd" / "lib" / "libxfdtd_core_amd.so"
    if not lib.exists():
        from xfdtd_core_amd import build as b

        b.build()
    prog = ROOT / "tests" / "cpp" / "build" / "sphere_scatter"
    if not prog.exists():
        from xfdtd_core_amd import build as b

        b.build_cpp_programs()


_ensure_built()


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:  # pragma: no cover
        return False


@pytest.fixture(scope="session")
def oracle_threads():
    return max(1, min(8, os.cpu_count() or 1))
