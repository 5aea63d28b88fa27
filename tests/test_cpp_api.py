"""The C++ scenario API (include/xfdtd_core_amd.hpp) as user code sees it.

* The reference README's program (README.md:73-122) and its simple_circuit.cpp
  example compile and link against this header and library with nothing
  changed but the #include lines (read from the reference tree where it is
  present; skipped elsewhere).  No copy of them is kept here.
* tests/cpp/sphere_scatter.cpp, the README scenario written as a user of the
  reference writes it plus monitors written by output(), visitors, the time
  axis and the constants, is built by the package build and links.
* On the GPU that program's output files (the movie's %05d.npy frames, a
  FieldTimeMonitor's (2, nt) file, a FieldMonitor's box) equal the oracle's
  arrays of the same scenario, with the reference's .npy header (<f8, C
  order), and the visitors see the reference's flags.
"""

import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path("/root/reference")
PROG = ROOT / "tests" / "cpp" / "build" / "sphere_scatter"
LIB_DIR = ROOT / "xfdtd_core_amd" / "lib"


def _compile_mapped(src_text: str, tmp_path: Path, name: str):
    """Maps every reference include to this library's header and compiles
    and links the program; returns the g++ result."""
    mapped = re.sub(r'#include [<"]xfdtd/[^>"]+[>"]', '#include "xfdtd_core_amd.hpp"', src_text)
    src = tmp_path / f"{name}.cpp"
    src.write_text(mapped)
    exe = tmp_path / name
    return subprocess.run(["g++", "-std=c++20", f"-I{ROOT / 'include'}", str(src), "-o", str(exe),
                           f"-L{LIB_DIR}", "-lxfdtd_core_amd"], capture_output=True, text=True)


@pytest.mark.skipif(not (REF / "README.md").exists(), reason="reference tree not present")
def test_reference_readme_program_compiles_and_links(tmp_path):
    txt = (REF / "README.md").read_text()
    progs = [b for b in re.findall(r"```cpp\n(.*?)```", txt, re.S) if "int main" in b]
    assert len(progs) == 1
    assert "addDefaultVisitor" in progs[0]
    r = _compile_mapped(progs[0], tmp_path, "readme")
    assert r.returncode == 0, r.stderr[-4000:]


@pytest.mark.skipif(not (REF / "test/src/examples/simple_circuit.cpp").exists(),
                    reason="reference tree not present")
def test_reference_simple_circuit_example_compiles_and_links(tmp_path):
    r = _compile_mapped((REF / "test/src/examples/simple_circuit.cpp").read_text(), tmp_path, "simple_circuit")
    assert r.returncode == 0, r.stderr[-4000:]


def test_scenario_program_is_built_and_linked():
    assert PROG.exists()
    out = subprocess.run(["ldd", str(PROG)], capture_output=True, text=True).stdout
    line = [ln for ln in out.splitlines() if "libxfdtd_core_amd.so" in ln]
    assert line and "not found" not in line[0], out
    # the header carries the reference's surface the program relies on
    hdr = (ROOT / "include" / "xfdtd_core_amd.hpp").read_text()
    for sym in ("addDefaultVisitor", "addVisitor", "SimulationFlagVisitor", "namespace constant",
                "calculationParam", "void output()", '"feild_monitor"', "calculateDt"):
        assert sym in hdr, sym


def _npy_header(path: Path) -> str:
    with open(path, "rb") as f:
        assert f.read(6) == b"\x93NUMPY"
        major, _ = f.read(2)
        assert major == 1
        n = int.from_bytes(f.read(2), "little")
        return f.read(n).decode()


def _oracle_spec():
    from xfdtd_core_amd import scenarios
    from xfdtd_core_amd import spec as S

    dl = 7.5e-3
    s = scenarios.dielectric_sphere()
    s.addMonitor(S.FieldTimeMonitor("ex_centre", S.Cube(S.Vector(0, 0, 0), S.Vector(dl, dl, dl)), "", "EX"))
    s.addMonitor(S.FieldMonitor(S.Cube(S.Vector(-0.175, -0.175, 0), S.Vector(0.35, 0.35, dl)), "EZ", "ez_xy", ""))
    return s.spec()


@pytest.mark.gpu
def test_scenario_program_outputs_match_oracle(tmp_path):
    from oracle import RefSim

    steps = 200
    print("[gpu] C++ scenario program: run", flush=True)
    r = subprocess.run([str(PROG), str(steps), str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    # the default visitor's lines on stderr, the counting visitor's on stdout
    for line in ("Simulation start", "Simulation update start", "Simulation update end", "Simulation end",
                 "Progress: ["):
        assert line in r.stderr, line
    lines = {ln.split(":", 1)[0]: ln.split(":", 1)[1] for ln in r.stdout.splitlines() if ":" in ln}
    flags = dict(re.findall(r"(\w+) (\d+)", lines["flags"]))
    assert flags["sim_start"] == flags["sim_end"] == flags["update_start"] == flags["update_end"] == "1"
    assert flags["last"] == flags["end"] == flags["current"] == str(steps)
    assert flags["monotonic"] == "1" and int(flags["next"]) >= 50
    tok = lines["dt"].split()  # dt, the nominal-dl dt, "c0:", C, "z0:", Z
    dt, dt_nominal, c0, z0 = float(tok[0]), float(tok[1]), float(tok[3]), float(tok[5])
    # the run's dt comes from the generated grid's smallest cell (0.35 m over
    # 47 cells), TimeParam::calculateDt(0.9, dl, dl, dl) from the nominal dl
    assert dt != dt_nominal and abs(dt / dt_nominal - 1) < 1e-2
    assert z0 == np.sqrt(4 * np.pi * 1e-7 / 8.854187817e-12)
    assert c0 == 1 / np.sqrt(8.854187817e-12 * 4 * np.pi * 1e-7)
    n_t, h0, h1, e0, e1 = (float(v) for v in lines["time"].split())
    # time_param.cpp:63-71: hTime = linspace(0.5 dt, (size + 0.5) dt, size), whose last sample is
    # (size + 0.5) dt
    assert n_t == steps and h0 == 0.5 * dt and h1 == (steps + 0.5) * dt
    assert e0 == dt and e1 == steps * dt

    print("[gpu] C++ scenario program: oracle", flush=True)
    ref = RefSim(_oracle_spec())
    ref.run(steps)
    assert dt == ref.dt
    frames = ref.monitor_records(0)
    assert len(frames) == steps // 10
    for q, want in enumerate(frames):
        f = tmp_path / "movie" / f"{10 * q:05d}.npy"
        hdr = _npy_header(f)
        assert "'descr': '<f8'" in hdr and "'fortran_order': False" in hdr, hdr
        got = np.load(f)
        assert got.shape == want.shape
        np.testing.assert_array_equal(got, want, err_msg=f.name)
    probe = np.load(tmp_path / "probe" / "ex_centre.npy")
    want = ref.monitor_records(1)[0]
    assert probe.shape == (2, steps) == want.shape
    np.testing.assert_array_equal(probe, want)
    plane = np.load(tmp_path / "plane" / "ez_xy.npy")
    want = ref.monitor_records(2)[0]
    assert plane.shape == want.shape and np.abs(want).max() > 0
    np.testing.assert_array_equal(plane, want)
    assert not list((tmp_path / "movie").glob("movie*.npy"))  # MovieMonitor::output writes nothing
