"""CRPA processing on the CPU: the small dense solver, the jammers' frequency offsets, the
hybrid's circular channels, the per-pass phasor fit, the blind run's tone separation and I/Q
snapshots, and the MVDR weights' defining properties, on synthetic port records with closed-form
answers; the receiver noise's draw order against GCC's own evaluation."""
import math
import shutil
import subprocess

import numpy as np
import pytest

from em_lab_amd import crpa as CR
from em_lab_amd.stdrandom import Mt19937, NormalDistribution


def test_solve_matches_lapack_and_zero_pivot():
    rng = np.random.default_rng(2)
    for n in (1, 2, 4, 7):
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 3 * np.eye(n)
        b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        np.testing.assert_allclose(CR.solve(A, b, n), np.linalg.solve(A, b), rtol=1e-12, atol=1e-12)
    # a vanishing pivot leaves its unknown at 0 (src/Crpa.cpp:43,52)
    x = CR.solve(np.array([[2.0, 0.0], [0.0, 0.0]]), np.array([4.0, 1.0]), 2)
    np.testing.assert_array_equal(x, [2.0, 0.0])


def _monitor(n_el=2, blind=False, wavelength=20.0, dt=0.5):
    m = CR.CrpaMonitor.__new__(CR.CrpaMonitor)
    m.setup = CR.CrpaSetup(satellite="sat", blind=blind, pass_periods=20.0, snapshot_periods=16.0,
                           snapshots_per_period=4,
                           elements=[CR.CrpaElement(f"x{i}", f"y{i}", (50.0 + 10 * i, 50.0, 30.0)) for i in range(n_el)])
    m.element_ports = [(2 * i, 2 * i + 1) for i in range(n_el)]
    m.center = np.mean([np.asarray(e.center, np.float32) for e in m.setup.elements], axis=0).astype(np.float32)
    m.wavelength, m.dt = wavelength, dt
    m.cell_size = 1e-3
    m.satellite = 0
    m.settle = []
    m.baseband = []
    return m


def test_emitter_offsets():
    m = _monitor(blind=True)
    m.satellite = 1
    m.setup.snapshot_periods = 40.0
    assert [m.emitter_offset(e) for e in range(5)] == [1 / 40, 0.0, -1 / 40, 2 / 40, -2 / 40]
    m.setup.blind = False
    assert m.emitter_offset(0) == 0.0


def test_hybrid_circular_channels():
    m = _monitor(n_el=1)
    # X = 1, Y = j: (X - jY) / sqrt 2 = sqrt 2 in the RHCP channel, nothing in the LHCP one
    np.testing.assert_allclose(m.channels([1.0, 1j]), [math.sqrt(2.0)])
    np.testing.assert_allclose(m.channels([1.0, 1j], lhcp=True), [0.0], atol=1e-15)
    m.element_ports = [(0, -1)]  # a missing probe reads 0
    np.testing.assert_allclose(m.channels([2.0]), [2.0 / math.sqrt(2.0)])


class _Rec:
    def __init__(self, v):
        self.v = v

    def read_series(self):
        out = np.zeros(self.v.shape + (2,), np.float32)
        out[..., 0] = self.v
        return out


def test_pass_fit_recovers_phasor():
    m = _monitor(n_el=1)
    m.response = [[], []]
    m.pass_ = 1
    w = 2 * math.pi / m.wavelength
    n = 2000
    t = (np.arange(n) + 1) * m.dt
    V = [0.7 * np.exp(1j * 0.4), 1.3 * np.exp(-1j * 2.1)]
    v = np.stack([np.real(Vp * np.exp(1j * w * t)) for Vp in V], axis=1)
    m.recorder = _Rec(v.astype(np.float32))
    m.finish_pass()
    np.testing.assert_allclose(m.response[1], V, rtol=1e-5, atol=1e-6)
    assert m.settle[1] < 1e-5


def test_blind_separation_and_snapshots():
    m = _monitor(n_el=1, blind=True)
    m.emitters = [CR.EmitterSetup("sat"), CR.EmitterSetup("jam")]
    lam = m.wavelength
    m.emitter_wavelength = [lam, lam / (1.0 + m.emitter_offset(1))]
    n = int(math.ceil((m.setup.pass_periods + m.setup.snapshot_periods + 4.5) * lam / m.dt))
    t = (np.arange(n) + 1) * m.dt
    ws = [2 * math.pi / wl for wl in m.emitter_wavelength]
    V = [[0.01 * np.exp(1j * 0.3), 0.02 * np.exp(1j * 1.0)], [0.5 * np.exp(-1j * 0.8), 0.4 * np.exp(1j * 2.4)]]
    v = np.zeros((n, 2))
    for e in range(2):
        for p in range(2):
            v[:, p] += np.real(V[e][p] * np.exp(1j * ws[e] * t))
    m.recorder = _Rec(v.astype(np.float32))
    m.finish_blind()
    for e in range(2):
        np.testing.assert_allclose(m.response[e], V[e], rtol=2e-4, atol=2e-6)
    assert len(m.baseband) == int(round(m.setup.snapshot_periods * 4))
    # each snapshot: the carrier component's phasor plus the jammer's, rotating at its offset
    # (the satellite is on the carrier: its baseband is steady)
    K = len(m.baseband)
    ta = (m.setup.pass_periods + 2.0) * lam
    tk = ta + np.arange(K) * lam / 4
    jam = m.channels(V[1])[0] * np.exp(1j * (ws[1] - ws[0]) * tk)
    got = np.array([b[0] for b in m.baseband])
    want = m.channels(V[0])[0] + jam
    # the 4-period low-pass passes a 1/16 offset tone with a small loss
    assert np.max(np.abs(got - want)) < 0.05 * np.max(np.abs(want))


def test_mvdr_properties():
    m = _monitor(n_el=4)
    m.setup.elements = [CR.CrpaElement(f"x{i}", f"y{i}", (50.0 + 10 * (i % 2), 50.0 + 10 * (i // 2), 30.0))
                        for i in range(4)]
    m.element_ports = [(2 * i, 2 * i + 1) for i in range(4)]
    m.center = np.mean([np.asarray(e.center, np.float32) for e in m.setup.elements], axis=0).astype(np.float32)
    m.emitters = [CR.EmitterSetup("sat", (55.0, 55.0, 130.0)), CR.EmitterSetup("j1", (160.0, 60.0, 40.0)),
                  CR.EmitterSetup("j2", (-40.0, 80.0, 45.0))]
    k = 2 * math.pi / m.wavelength

    def response(pos):  # plane-wave RHCP arrival on each element's two probes
        u = np.asarray(pos, np.float64) - m.center
        u /= np.linalg.norm(u)
        out = []
        for el in m.setup.elements:
            ph = np.exp(1j * k * float((np.asarray(el.center) - m.center) @ u))
            out += [ph, 1j * ph]
        return out

    m.response = [response(e.position) for e in m.emitters]
    m.settle = [0.0, 0.0, 0.0]
    m.compute_weights()
    r = m.result
    assert r.valid
    a = np.array([np.exp(1j * k * float((np.asarray(el.center, np.float32) - m.center) @
                                        ((np.asarray(m.emitters[0].position, np.float32) - m.center) /
                                         np.float32(np.linalg.norm(np.asarray(m.emitters[0].position, np.float32) - m.center)))))
                  for el in m.setup.elements])
    assert CR.dot_h(r.weights, a) == pytest.approx(1.0, abs=1e-9)  # distortionless toward the satellite
    jam_out = [e.output_db for e in r.emitters if not e.satellite]
    jam_in = [e.input_db for e in r.emitters if not e.satellite]
    assert all(o < i - 30.0 for o, i in zip(jam_out, jam_in))       # nulls on both jammers
    assert r.sinr_mvdr_db > r.sinr_steering_db + 20.0 and r.sinr_mvdr_db > r.sinr_single_db
    assert r.map_db.size == 61 * 61 and np.nanmax(r.steering_map_db) == pytest.approx(0.0, abs=1e-6)


@pytest.mark.skipif(shutil.which("g++") is None, reason="no C++ compiler")
def test_blind_noise_draw_order_matches_gcc(tmp_path):
    """cd(gauss(rng), gauss(rng)) (src/Crpa.cpp:443): the two draws' order is the compiler's; GCC
    evaluates the imaginary part's first, which compute_weights follows."""
    src = tmp_path / "ev.cpp"
    src.write_text("""#include <random>
#include <complex>
#include <cstdio>
int main() {
  std::mt19937 rng(3u); std::normal_distribution<double> g(0.0, 0.7071067811865476);
  for (int i = 0; i < 5; ++i) { std::complex<double> c(g(rng), g(rng)); std::printf("%.17g %.17g\\n", c.real(), c.imag()); }
}
""")
    exe = tmp_path / "ev"
    subprocess.run(["g++", "-O2", str(src), "-o", str(exe)], check=True)
    want = [float(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    e, g = Mt19937(3), NormalDistribution(0.0, 0.7071067811865476)
    got = []
    for _ in range(5):
        im = g(e)
        re = g(e)
        got += [re, im]
    assert got == want
