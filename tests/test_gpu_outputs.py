"""Output files of the GPU path read back and checked against the oracle.

The reference writes its results as .npy files (xt::dump_npy, `<f8` or
`<c16`, C order, monitor.cpp:42-54): movie frames `%05d.npy`
(movie_monitor.cpp:26-34), time series as the (2, nt) stack of time and data
(time_monitor.cpp:11-21), the frequency-domain far field as
`<f>GHz_{a_theta,f_phi,a_phi,f_theta}.npy` and the radiated power as the
(2, nf) stack `power.npy` (nffft_frequency_domain.cpp:65-186).  Each file the
device run writes is read back, its header checked, and its array compared
with the oracle's array of the same scenario (FP64: bitwise for the fields,
1e-12 for the far-field sums).
"""

import numpy as np
import pytest

from oracle import RefSim, rel_l2
from xfdtd_core_amd import native
from xfdtd_core_amd import spec as S

pytestmark = pytest.mark.gpu

FREQS = [2.0e9, 3.5e9]
STEPS = 160


def _header(path):
    with open(path, "rb") as f:
        assert f.read(6) == b"\x93NUMPY"
        assert f.read(2)[0] == 1
        n = int.from_bytes(f.read(2), "little")
        return f.read(n).decode()


def _scenario(out):
    dl = 5e-3
    s = S.Simulation(dl, dl, dl, 0.9)
    s.addObject(S.Object("domain", S.Cube(S.Vector(-0.06, -0.06, -0.06), S.Vector(0.12, 0.12, 0.12)),
                         S.Material.createAir()))
    s.addObject(S.Object("ball", S.Sphere(S.Vector(0.0, 0.0, 0.0), 0.03),
                         S.Material("b", S.ElectroMagneticProperty(3.0, 2.0, 0.01, 0.0))))
    tau = 20 * dl / 6e8
    s.addWaveformSource(S.TFSF3D(9, 9, 9, 0.3, 0.5, 0.2, S.Waveform.gaussian(tau, 4.5 * tau)))
    for d in ("XN", "XP", "YN", "YP", "ZN", "ZP"):
        s.addBoundary(S.PML(6, d))
    s.addMonitor(S.MovieMonitor(S.FieldMonitor(S.Cube(S.Vector(-0.06, 0.0, -0.06), S.Vector(0.12, dl, 0.12)), "EZ"),
                                20, "movie", str(out / "movie") if out else ""))
    s.addMonitor(S.FieldTimeMonitor("hy_probe", S.Cube(S.Vector(0.01, 0.0, 0.02), S.Vector(dl, dl, dl)),
                                    str(out / "probe") if out else "", "HY"))
    s.addMonitor(S.VoltageMonitor("v_gap", S.Cube(S.Vector(-0.02, 0.0, -0.01), S.Vector(0.0, 0.0, 0.02)), "ZP",
                                  str(out / "volt") if out else ""))
    s.addMonitor(S.FieldMonitor(S.Cube(S.Vector(-0.06, -0.06, 0.0), S.Vector(0.12, 0.12, dl)), "EX", "ex_final",
                                str(out / "final") if out else ""))
    s.addNF2FF(S.NFFFTFrequencyDomain(8, 8, 8, FREQS, str(out / "nf2ff") if out else ""))
    return s.spec()


def test_output_files_match_oracle(tmp_path):
    print("[gpu] output files: device", flush=True)
    dev = native.Simulation(_scenario(tmp_path), "f64")
    dev.run(STEPS)
    for m in range(dev.monitor_count()):
        dev.monitor_output(m)
    theta = np.linspace(0.0, np.pi, 37)
    dev.nf2ff_process_far_field(0, theta, 0.7, "ff_phi07")
    dev.nf2ff_output_power(0)
    print("[gpu] output files: oracle", flush=True)
    ref = RefSim(_scenario(None))
    ref.run(STEPS)

    # movie frames %05d.npy of the record count
    frames = ref.monitor_records(0)
    assert len(frames) == STEPS // 20
    for q, want in enumerate(frames):
        f = tmp_path / "movie" / f"{20 * q:05d}.npy"
        h = _header(f)
        assert "'descr': '<f8'" in h and "'fortran_order': False" in h and f"{want.shape}" in h, h
        np.testing.assert_array_equal(np.load(f), want, err_msg=f.name)
    assert sorted(p.name for p in (tmp_path / "movie").iterdir()) == [f"{20 * q:05d}.npy" for q in range(len(frames))]
    # field-time and voltage monitors: (2, nt) = stack(time, data)
    for m, path in ((1, tmp_path / "probe" / "hy_probe.npy"), (2, tmp_path / "volt" / "v_gap.npy")):
        want = ref.monitor_records(m)[0]
        got = np.load(path)
        h = _header(path)
        assert f"'shape': (2, {STEPS})" in h and "'descr': '<f8'" in h, h
        np.testing.assert_array_equal(got, want)
        assert np.abs(got[1]).max() > 0
    # a FieldMonitor on its own: the box of the final fields
    want = ref.monitor_records(3)[0]
    got = np.load(tmp_path / "final" / "ex_final.npy")
    assert got.shape == want.shape and np.abs(want).max() > 0
    np.testing.assert_array_equal(got, want)
    # far-field potentials <f>GHz_*.npy (complex128) and power.npy (2, nf)
    for fi, fq in enumerate(FREQS):
        tag = f"{fq / 1e9:.2f}GHz"
        for name, a_pot, th_comp in (("a_theta", True, True), ("f_phi", False, False), ("a_phi", True, False),
                                     ("f_theta", False, True)):
            path = tmp_path / "nf2ff" / "ff_phi07" / f"{tag}_{name}.npy"
            h = _header(path)
            assert "'descr': '<c16'" in h and f"'shape': ({theta.size},)" in h, h
            got = np.load(path)
            want = ref.nf2ff_potential(0, fi, a_pot, th_comp, theta, 0.7)
            assert rel_l2(got, want) <= 1e-12, (tag, name, rel_l2(got, want))
    power = np.load(tmp_path / "nf2ff" / "power.npy")
    assert power.shape == (2, len(FREQS))
    np.testing.assert_array_equal(power[0], FREQS)
    for fi in range(len(FREQS)):
        want = ref.nf2ff_power(0, fi)
        assert abs(power[1, fi] - want) <= 1e-12 * abs(want), (fi, power[1, fi], want)
    dev.close()
