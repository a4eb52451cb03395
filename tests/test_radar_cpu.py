"""Radar host side and its oracle on the CPU: the standard-library noise generators against the
C++ standard's known answer and libstdc++ itself, the imaging oracle's closed forms, the host
processing (taper, parabola refinement, pulse compression, detection) on synthetic records."""
import math
import shutil
import subprocess

import numpy as np
import pytest

from em_lab_amd import radar as R
from em_lab_amd.stdrandom import Mt19937, NormalDistribution
from oracle import radar as OR


def test_mt19937_known_answer():
    # ISO C++ [rand.predef]: the 10000th draw of a default-constructed std::mt19937
    e = Mt19937()
    for _ in range(9999):
        e()
    assert e() == 4123659995


@pytest.mark.skipif(shutil.which("g++") is None, reason="no C++ compiler to run libstdc++'s generator")
def test_normal_distribution_matches_libstdcxx(tmp_path):
    src = tmp_path / "nd.cpp"
    src.write_text("""#include <random>
#include <cstdio>
int main() {
  for (unsigned seed : {20251105u, 20251106u, 1u}) {
    std::mt19937 rng(seed); std::normal_distribution<double> g(0.0, 0.7071067811865476);
    for (int i = 0; i < 257; ++i) std::printf("%.17g\\n", g(rng));
  }
}
""")
    exe = tmp_path / "nd"
    subprocess.run(["g++", "-O1", str(src), "-o", str(exe)], check=True)
    want = [float(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    got = []
    for seed in (20251105, 20251106, 1):
        e, g = Mt19937(seed), NormalDistribution(0.0, 0.7071067811865476)
        got += [g(e) for _ in range(257)]
    assert got == want


AXES = np.array([[0, -1, 0], [0, 0, 1], [1, 0, 0]], np.float32)  # boresight +x: right = -y, up = +z


def test_oracle_direction_axes():
    np.testing.assert_allclose(OR.direction(AXES, 0, 0), AXES[2], atol=1e-7)
    np.testing.assert_allclose(OR.direction(AXES, 90, 0), AXES[0], atol=1e-6)
    np.testing.assert_allclose(OR.direction(AXES, 0, 90), AXES[1], atol=1e-6)
    d = OR.direction(AXES, 30, 20)
    assert abs(np.linalg.norm(d) - 1.0) < 1e-6
    assert d @ AXES[0] == pytest.approx(math.sin(math.radians(30)) * math.cos(math.radians(20)), abs=1e-6)


def test_oracle_tx_delay_single_transmitter_is_distance():
    rng = np.random.default_rng(3)
    tx = np.array([10.0, 20.0, 20.0], np.float32)
    for _ in range(20):
        P = tx + rng.uniform([5, -10, -10], [60, 10, 10]).astype(np.float32)
        d = OR.tx_delay(P, tx, AXES, [tx], [rng.uniform(0.2, 3.0)], 20.0)
        assert d == pytest.approx(float(np.linalg.norm(P - tx)), rel=1e-6)


def _setup(nr=6, nel=5, naz=5, sample_dt=2.0, dt=0.5):
    return dict(tx=(10.0, 20.0, 20.0), axes=AXES, interior_min=(4.0, 4.0, 4.0), interior_max=(70.0, 36.0, 36.0),
                wavelength=16.0, dt=dt, sample_dt=sample_dt, r0=8.0, dr=5.0, nr=nr, az0=-30.0, daz=15.0, naz=naz,
                el0=-30.0, del_=15.0, nel=nel, **{"del": 15.0})


def test_oracle_single_element_constant_echo():
    st = _setup()
    m = 30  # 60 time units: the farthest points' round trips run past the record
    c = np.complex64(0.3 - 0.7j)
    z = np.full((m, 1), c, np.complex64)
    p = OR.back_project(st, z, [(10.0, 21.0, 20.0)], [1.7], [st["tx"]], [1.0])
    inside = p >= 0.0
    assert inside.any() and (~inside).any()  # some points fall outside the interior
    expect = (1.7 * abs(complex(c))) ** 2
    vals = p[inside]
    # inside the record: w^2 |c|^2 (the float interpolation of a constant rounds); past its end: 0
    near = np.abs(vals - expect) < 1e-6 * expect
    assert np.all(near | (vals == 0.0))
    assert np.any(near) and np.any(vals == 0.0)


def test_oracle_coherent_sum_at_the_chosen_point():
    st = _setup(nr=1, nel=1, naz=1, sample_dt=0.5)
    st.update(r0=30.0, az0=0.0, el0=0.0)
    tx = np.asarray(st["tx"], np.float32)
    P = tx + np.float32(30.0) * OR.direction(AXES, 0, 0)
    els = np.array([[10.0, 14.0, 20.0], [10.0, 26.0, 20.0]], np.float32)
    w = 2 * math.pi / 16.0
    sdt = st["sample_dt"]
    m = np.arange(256)
    z = np.zeros((256, 2), np.complex64)
    for i in range(2):
        # e^{-j w t} sampled at the decimated delays t = m sdt - dt / 2: the phase that the
        # back-projection's e^{+j w t} cancels at any delay, hence at the chosen point
        z[:, i] = np.exp(-1j * w * (m * sdt - 0.5 * st["dt"]))
    p = OR.back_project(st, z, els, [0.6, 0.9], [tx], [1.0])[0, 0, 0]
    # both terms in phase: (0.6 + 0.9)^2, less what linear interpolation of e^{-j w t} between
    # samples w sdt apart takes off (the chord's midpoint is cos(w sdt / 2) of the arc)
    assert 1.5 ** 2 * math.cos(w * sdt / 2) ** 2 * (1 - 1e-6) <= p <= 1.5 ** 2 * (1 + 1e-6)


def test_taper_hamming_values():
    mon = R.RadarMonitor.__new__(R.RadarMonitor)
    # 4 x 1 elements along the azimuth axis (-y): count 4, window over 6 points without its ends
    pos = np.array([[10, 20 - 10 * k, 20] for k in range(4)], np.float32)
    coords = np.round((pos @ AXES[0]) * 4) / 4
    uniq = np.unique(coords)
    w = [0.54 - 0.46 * math.cos(2 * math.pi * (np.searchsorted(uniq, c) + 1) / 5) for c in coords]
    assert w == pytest.approx([0.54 - 0.46 * math.cos(2 * math.pi * k / 5) for k in (1, 2, 3, 4)])
    assert mon is not None


def test_vertex_offset_parabola():
    for x0 in (-0.4, -0.1, 0.0, 0.25, 0.45):
        f = [-(t - x0) ** 2 + 3.0 for t in (-1, 0, 1)]
        assert R.vertex_offset(*f) == pytest.approx(x0, abs=1e-6)
    assert R.vertex_offset(1.0, 1.0, 1.0) == 0.0
    assert R.vertex_offset(0.0, 0.1, 5.0) == -0.5 and R.vertex_offset(5.0, 0.1, 0.0) == 0.5  # clamped


def test_boresight_frame():
    r, u = R.boresight_frame((1, 0, 0))
    np.testing.assert_array_equal(r, [0, -1, 0])
    np.testing.assert_array_equal(u, [0, 0, 1])
    r, u = R.boresight_frame((0, 0, -1))
    np.testing.assert_array_equal(r, [1, 0, 0])
    np.testing.assert_array_equal(u, [0, 1, 0])


def _synthetic_monitor(elements, tx, wavelength=16.0, dt=0.5, n=1024):
    """A monitor with a transmitted Gaussian pulse, a filter matched to it and no calibration
    response, as computeImage sees them after buildTemplate."""
    mon = R.RadarMonitor.__new__(R.RadarMonitor)
    mon.elements = [None] * len(elements)
    mon.element_pos = np.asarray(elements, np.float32)
    mon.weights = np.ones(len(elements))
    mon.axes = AXES
    mon.tx = np.asarray(tx, np.float32)
    mon.tx_pos = mon.tx[None, :]
    mon.tx_amp = np.ones(1)
    mon.wavelength, mon.dt = wavelength, dt
    mon.setup = R.RadarSetup(range_min=10.0, range_max=60.0, fov_az_deg=30.0, fov_el_deg=30.0)
    mon.cal = np.zeros((n, len(elements) + 3), np.float32)
    mon.cal_steps = n
    F = 1
    while F < 2 * n:
        F <<= 1
    mon.fft_size = F
    s = R.E.SourceDesc(wavelength=wavelength, waveform=R.E.WAVE_PULSE, ramp_periods=0.5)
    pulse = lambda t: R.source_signal(s, t)  # noqa: E731
    drive = np.zeros(F)
    drive[:n] = [pulse((k + 0.5) * dt) for k in range(n)]
    S = np.fft.fft(drive)
    w = 2 * math.pi / wavelength
    om = np.arange(F) * 2 * math.pi / (F * dt)
    mon.filter = np.where((np.arange(F) >= 1) & (np.arange(F) < F // 2),
                          np.conj(S) * np.exp(-(((om - w) / (0.3 * w)) ** 2)), 0)
    mon.peak_tx = 1.0
    mon.template_end = 0.0
    mon.extent = 0.0
    mon.system_delay = 0.0
    mon.targets, mon.target_nearest = [], []
    mon.interior_min = np.zeros(3, np.float32)
    mon.interior_max = np.full(3, 1e4, np.float32)
    return mon, pulse


def test_compression_and_imaging_find_a_point_scatterer():
    """Echoes of the drive pulse from a point at a chosen range and angle, compressed with the
    matched filter and back-projected (oracle): the strongest image point is that point's bin."""
    els = [(10.0, 20.0 + dy, 20.0 + dz) for dy in (-6.0, 6.0) for dz in (-6.0, 6.0)]
    tx = (10.0, 20.0, 20.0)
    mon, pulse = _synthetic_monitor(els, tx)
    rng_true, az_true, el_true = 35.0, 12.0, -6.0
    P0 = np.asarray(tx, np.float32) + np.float32(rng_true) * OR.direction(AXES, az_true, el_true)
    n = mon.cal_steps
    meas = np.zeros_like(mon.cal)
    for i, e in enumerate(els):
        tau = np.linalg.norm(P0 - np.asarray(tx)) + np.linalg.norm(P0 - np.asarray(e, np.float32))
        meas[:, i] = [pulse((k + 1) * mon.dt - tau) for k in range(n)]
    z, D, M = mon.compressed_echoes(meas, n)
    st = dict(tx=mon.tx, axes=AXES, interior_min=mon.interior_min, interior_max=mon.interior_max,
              wavelength=mon.wavelength, dt=mon.dt, sample_dt=D * mon.dt, r0=31.0, dr=1.0, nr=9, az0=6.0, daz=3.0,
              naz=5, el0=-12.0, nel=5, **{"del": 3.0})
    p = OR.back_project(st, z, mon.element_pos, mon.weights, mon.tx_pos, mon.tx_amp)
    ir, ie, ia = np.unravel_index(int(np.argmax(p)), p.shape)
    assert 31.0 + ir * 1.0 == pytest.approx(rng_true, abs=1.0)
    assert 6.0 + ia * 3.0 == pytest.approx(az_true, abs=3.0)
    assert -12.0 + ie * 3.0 == pytest.approx(el_true, abs=3.0)


def test_detection_on_a_planted_peak():
    mon, _ = _synthetic_monitor([(10.0, 20.0, 20.0)], (10.0, 20.0, 20.0))
    mon.setup = R.RadarSetup(threshold_db=13.0, dynamic_range_db=25.0)
    img = R.RadarImage(nr=40, nel=21, naz=21, r0=10.0, dr=2.0, az0=-30.0, daz=3.0, el0=-30.0, del_=3.0)
    r = np.arange(40)[:, None, None]
    e = np.arange(21)[None, :, None]
    a = np.arange(21)[None, None, :]
    rng = np.random.default_rng(5)
    power = 1e-4 * rng.random((40, 21, 21)) + np.exp(-((r - 17.3) ** 2 / 2 + (e - 12.6) ** 2 / 2 + (a - 7.2) ** 2 / 2))
    meas = np.zeros_like(mon.cal)
    out = mon.finish_image(img, power, meas, mon.cal_steps, True)
    assert out.valid and len(out.detections) >= 1
    d = out.detections[0]
    assert d.rel_db == pytest.approx(0.0, abs=0.5)
    assert d.range == pytest.approx(10.0 + 17.3 * 2.0, abs=0.3)
    assert d.el_deg == pytest.approx(-30.0 + 12.6 * 3.0, abs=0.5)
    assert d.az_deg == pytest.approx(-30.0 + 7.2 * 3.0, abs=0.5)
    assert d.snr_db > 13.0
    # display slices: the B-scope is the maximum over elevation, the A-scope runs through the peak
    np.testing.assert_array_equal(out.bscope, out.power_db.max(axis=1))
    assert int(np.argmax(out.ascope)) == 17
