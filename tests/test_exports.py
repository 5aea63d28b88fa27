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


def test_ctypes_descriptors_match_the_c_header(tmp_path):
    """The Python descriptor structs have the C header's size and field
    offsets (compiled with the host C compiler)."""
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {
        "xfd_cpml_desc": native.CpmlDesc, "xfd_tfsf_desc": native.TfsfDesc, "xfd_lumped_desc": native.LumpedDesc,
        "xfd_gather_desc": native.GatherDesc, "xfd_desc": native.Desc,
    }
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "xfdtd_core_amd.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} __sizeof__ %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, py in structs.items():
        assert got[(cname, "__sizeof__")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
