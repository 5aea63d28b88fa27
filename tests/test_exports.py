"""The C-ABI library loads and exports every function the headers declare.

No GPU is needed: loading the library and resolving symbols does not call
into HIP.
"""

import ctypes
import re
from pathlib import Path

import pytest

from xfdtd_core_amd import native

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions():
    names = []
    for h in HEADERS:
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(xfd_[a-z0-9_]+)\s*\(", text):
            names.append(m.group(1))
    return sorted(set(names))


def test_headers_declare_functions():
    names = declared_functions()
    assert "xfd_create" in names and "xfd_step" in names and "xfd_sim_run" in names
    assert len(names) > 20


@pytest.mark.parametrize("name", declared_functions())
def test_symbol_exported(name):
    lib = ctypes.CDLL(str(native.library_path()))
    assert hasattr(lib, name), name


def test_version_string():
    L = native.load_library()
    assert b"gfx950" in L.xfd_version()
