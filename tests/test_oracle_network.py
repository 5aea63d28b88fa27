"""Pins the oracle's S-parameter post-process (util/dft.h:9-26,
port.cpp:44-130, network.cpp:47-84) with closed forms:

* dft: an impulse at sample j0 transforms to dt exp(-2 pi i f ((j0+1) dt +
  shift)); at the FFT bin frequencies k / (N dt) the transform equals
  dt exp(-2 pi i k / N) FFT(x)_k (numpy's FFT, an independent routine).
* power waves: an open port (no current) reflects S11 = 1, a shorted one
  (no voltage) S11 = -conj(z) / z; a port whose current record is v / z
  (a matched line, sampled at the voltage's times) reflects
  S11 = -i tan(theta / 2), theta = 2 pi f 1.5 dt, because the current's
  transform is taken 1.5 dt earlier -- the sign and size of that shift.
* Network: S(j, i) = b_j / a_i keyed 10 j + i for the source ports only,
  on the oracle's own monitor records of a two-port cavity run; the .npy
  files carry the same values.
"""

import numpy as np
import pytest

import eigenmode as em
from oracle import RefSim, port_waves, series_dft
from xfdtd_core_amd import spec as S

DT = 1.7e-12
FREQS = np.array([1e9, 3.3e9, 7.5e9, 2.1e10])


@pytest.mark.parametrize("shift", [0.0, -1.5 * DT, 0.4 * DT])
def test_dft_of_impulse(shift):
    n, j0 = 40, 13
    x = np.zeros(n)
    x[j0] = 2.5
    got = series_dft(x, DT, FREQS, shift)
    want = 2.5 * DT * np.exp(-2j * np.pi * FREQS * ((j0 + 1) * DT + shift))
    np.testing.assert_allclose(got, want, rtol=1e-13)


def test_dft_matches_fft_bins():
    rng = np.random.default_rng(3)
    n = 64
    x = rng.standard_normal(n)
    k = np.arange(n // 2)
    f = k / (n * DT)
    got = series_dft(x, DT, f)
    want = DT * np.exp(-2j * np.pi * k / n) * np.fft.fft(x)[: n // 2]
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-11 * np.abs(want).max())


def _gauss(n=300):
    """A pulse whose spectrum is well above roundoff at every test frequency."""
    t = (np.arange(n) + 1) * DT
    tau = 10 * DT
    return np.exp(-(((t - 4 * tau) / tau) ** 2))


@pytest.mark.parametrize("z", [50.0, 75 + 20j])
def test_open_and_short_ports(z):
    v = _gauss()
    a, b = port_waves(np.zeros_like(v), v, DT, FREQS, z)
    np.testing.assert_allclose(b / a, np.ones(FREQS.size), rtol=1e-14)
    k = np.sqrt(complex(z).real)
    np.testing.assert_allclose(a, 0.5 * series_dft(v, DT, FREQS) / k, rtol=1e-14)
    a, b = port_waves(v, np.zeros_like(v), DT, FREQS, z)
    np.testing.assert_allclose(b / a, -np.conj(z) / z * np.ones(FREQS.size), rtol=1e-13)


def test_matched_line_reflects_the_current_shift():
    z = 50.0
    v = _gauss()
    a, b = port_waves(v / z, v, DT, FREQS, z)
    theta = 2 * np.pi * FREQS * 1.5 * DT
    np.testing.assert_allclose(b / a, -1j * np.tan(theta / 2), rtol=1e-12, atol=1e-14)


N = (24, 20, 16)
DL = (1e-3, 1.2e-3, 0.9e-3)
STEPS = 80


def _cube(o, e):
    return S.Cube(S.Vector(*[o[a] * DL[a] for a in range(3)]), S.Vector(*[(e[a] - o[a]) * DL[a] for a in range(3)]))


def two_port_cavity(out_dir=""):
    """A PEC cavity holding an Ez eigenmode with two ports (V/I monitor
    pairs along z), port 1 a source, port 2 not."""
    sp = em.cavity_spec(N, DL)
    v1 = S.VoltageMonitor("v1", _cube((6, 5, 3), (8, 7, 12)), "ZP")
    c1 = S.CurrentMonitor("c1", _cube((6, 5, 3), (8, 7, 9)), "ZP")
    v2 = S.VoltageMonitor("v2", _cube((15, 12, 3), (17, 14, 12)), "ZP")
    c2 = S.CurrentMonitor("c2", _cube((15, 12, 3), (17, 14, 9)), "ZP")
    nw = S.Network([S.Port(1, True, 50.0, c1, v1), S.Port(2, False, 50.0, c2, v2)], list(FREQS), out_dir)
    return sp + "".join(m.line() + "\n" for m in (v1, c1, v2, c2)) + nw.line() + "\n"


def test_network_keys_and_files(tmp_path):
    out = tmp_path / "net"
    ref = RefSim(two_port_cavity(str(out)))
    ref.init(STEPS)
    e0, _, _ = em.mode("EZ", N)
    ref.write_field("EZ", e0)
    ref.loop()
    recs = [ref.monitor_records(m)[0][1] for m in range(4)]  # v1 c1 v2 c2
    s = ref.network_s(0, (11, 21), FREQS.size)
    a1, b1 = port_waves(recs[1], recs[0], ref.dt, FREQS, 50.0)
    a2, b2 = port_waves(recs[3], recs[2], ref.dt, FREQS, 50.0)
    assert np.abs(b1).max() > 0 and np.abs(b2).max() > 0
    # numpy's complex division rounds differently from libstdc++'s by an ulp
    np.testing.assert_allclose(s[11], b1 / a1, rtol=1e-15)
    np.testing.assert_allclose(s[21], b2 / a1, rtol=1e-15)
    with pytest.raises(RuntimeError):  # port 2 is no source: no S(j, 2)
        ref.network_s(0, (12,), FREQS.size)
    np.testing.assert_array_equal(np.load(out / "s11.npy"), s[11])
    np.testing.assert_array_equal(np.load(out / "s21.npy"), s[21])
    np.testing.assert_array_equal(np.load(out / "frequencies.npy"), FREQS)
    assert np.load(out / "s21.npy").dtype == np.complex128
