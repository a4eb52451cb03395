"""Host-side pieces of the learned screening path and the batched scene generator: feature
encoding, the Gamma <-> Z map, the -10 dB bandwidth measure (against brute force), seeded design
generation, and the (theta, phi) direction grid."""
import math

import numpy as np
import pytest

from em_lab_amd import surrogate as S
from em_lab_amd.scene import (ARRAY_PML, draw_array_design, draw_array_designs, draw_array_geometry, draw_array_scene, spheres_array,
                              Xoshiro256ss)


def test_gamma_z_round_trip_and_s11():
    rng = np.random.default_rng(0)
    z = rng.uniform(1, 300, 100) + 1j * rng.uniform(-900, 300, 100)
    g = S.z_to_gamma(z)
    assert np.all(np.abs(g) < 1.0)  # passive loads map inside the unit circle
    assert np.allclose(S.gamma_to_z(g), z, rtol=1e-12)
    assert S.z_to_gamma(np.array([50.0 + 0j]))[0] == 0
    assert S.s11_db(np.array([0.1]))[0] == pytest.approx(-20.0)


@pytest.mark.parametrize("power", [0.0, 0.5, 0.25])
def test_companded_targets_invert(power):
    """targets_from_z / gamma_from_output are inverse at every companding power, keep the phase,
    map |Gamma| to |Gamma|^(1 - power) and Gamma = 0 to 0."""
    rng = np.random.default_rng(3)
    z = rng.uniform(1, 400, (5, S.N_BINS)) + 1j * rng.uniform(-600, 600, (5, S.N_BINS))
    z[0, 0] = S.R0  # Gamma exactly 0
    t = S.targets_from_z(z, power)
    assert t.shape == (5, 2 * S.N_BINS) and t.dtype == np.float32
    g = S.z_to_gamma(z)
    o = t[:, :S.N_BINS] + 1j * t[:, S.N_BINS:]
    assert o[0, 0] == 0
    m = np.abs(g) > 1e-6
    assert np.allclose(np.abs(o[m]), np.abs(g[m]) ** (1.0 - power), rtol=1e-5)
    assert np.allclose(np.angle(o[m]), np.angle(g[m]), atol=1e-5)
    back = S.gamma_from_output(t, power)
    assert np.allclose(back, g, rtol=2e-5, atol=1e-6)
    assert back[0, 0] == 0


def test_bandwidth_matches_brute_force():
    rng = np.random.default_rng(1)
    s = rng.uniform(-20, 0, (200, 64))
    nu = S.sweep_nu()
    bw = S.bandwidth_10db(s, nu)
    for i in range(200):
        best = run = 0
        for v in s[i]:
            run = run + 1 if v < -10 else 0
            best = max(best, run)
        assert bw[i] == pytest.approx(best * (nu[1] - nu[0]))
    assert nu[0] == pytest.approx(0.5) and nu[-1] == pytest.approx(1.5)


def test_features_encode_geometry_with_mask():
    g = draw_array_geometry(Xoshiro256ss(5))
    f = S.features(g)
    assert f.shape == (18,)
    for i in range(6):
        if i < g.elements:
            assert f[3 * i] == pytest.approx(g.lengths[i] / 16.0)
            assert f[3 * i + 2] == pytest.approx(math.log10(g.radii[i] / 16.0) + 3.5)
            assert f[3 * i] > 0
        else:
            assert not np.any(f[3 * i:3 * i + 3])


def test_design_generator_ranges_and_determinism():
    sizes, counts = [], []
    for i in range(300):
        g, size = draw_array_design(i)
        assert (g, size) == draw_array_design(i)  # seeded: same design every time
        assert 1 <= g.elements <= 6 and g.lengths[0] % 2 == 1 and g.lengths[0] >= 5
        assert all(1 <= s for s in g.spacings[1:]) and g.spacings[0] == 0
        assert all(0.001 * 16 <= r <= 0.01 * 16 * 1.0001 for r in g.radii)
        assert all(v % 8 == 0 for v in size)
        assert 48 <= size[0] <= 128 and 48 <= size[1] <= 128 and 96 <= size[2] <= 160
        assert size[0] >= g.span() + 2 * ARRAY_PML + 8
        counts.append(g.elements)
    assert set(counts) == {1, 2, 3, 4, 5, 6}
    g, sc = draw_array_scene(7)
    assert len(sc.ports) == 1 and len(sc.sources) == 1 and sc.sweep.points == 64
    assert len(sc.subcell.wire_terms) > 0 and not np.any(sc.sigma)


@pytest.mark.parametrize("start,count,seed", [(0, 3000, S.CANDIDATE_SEED), (123456, 1000, 7), (5, 1, 1)])
def test_lane_stepped_generator_equals_scalar(start, count, seed):
    """draw_array_designs (all xoshiro256** streams stepped together in uint64 lanes) gives the
    scalar generator's designs exactly, and features_many its features."""
    geoms, sizes = draw_array_designs(start, count, seed)
    for k in range(count):
        g, size = draw_array_design(start + k, seed)
        assert (geoms[k].lengths, geoms[k].spacings, geoms[k].radii, sizes[k]) == (g.lengths, g.spacings, g.radii, size)
    ref = np.stack([S.features(g) for g in geoms])
    assert np.array_equal(S.features_many(geoms), ref)


def test_spheres_array_scene():
    sc = spheres_array(96)
    assert len(sc.ports) == 16 and len(sc.sources) == 16
    texels = {p.texel for p in sc.ports}
    assert len(texels) == 16
    xs = sorted({t[0] for t in texels})
    assert len(xs) == 4 and all(b - a == 10 for a, b in zip(xs, xs[1:]))  # half-wave spacing at lambda 20


def test_theta_phi_grid():
    from em_lab_amd.emfdtd import FarFieldMonitor
    d = FarFieldMonitor.theta_phi_grid(91, 181)
    assert d.shape == (91 * 181, 4)
    assert np.allclose(np.linalg.norm(d[:, :3], axis=1), 1.0, atol=1e-6)
    assert np.allclose(d[0, :3], [0, 0, 1], atol=1e-6) and np.allclose(d[-1, :3], [0, 0, -1], atol=1e-6)


def test_report_csv_formats(tmp_path):
    """The sweep and pattern CSVs keep EM_Lab's columns and C++ stream number formatting."""
    from em_lab_amd import emfdtd as E
    pts = np.array([[20.0, 50.0, -3.25, -20.5, 1.2], [10.0, 73.125, 42.5, -3.1, 5.75]])
    sweep = dict(ports=[dict(points=pts)])
    p = tmp_path / "sweep.csv"
    assert E.write_sweep_csv(p, ["dipole"], sweep, 1.0e-3)
    lines = p.read_text().splitlines()
    assert lines[0] == "frequency_Hz,wavelength_cells,dipole_R_ohm,dipole_X_ohm,dipole_S11_dB,dipole_VSWR"
    assert lines[1] == "1.49896e+10,20,50,-3.25,-20.5,1.2"
    assert lines[2] == "2.99792e+10,10,73.125,42.5,-3.1,5.75"
    pat = dict(valid=True, directivity_dbi=np.arange(540, dtype=np.float32) / 100.0)
    q = tmp_path / "pattern.csv"
    assert E.write_pattern_csv(q, pat)
    rows = q.read_text().splitlines()
    assert rows[0] == "angle_deg,xy_dBi,xz_dBi,yz_dBi" and len(rows) == 181
    assert rows[2] == "2,0.01,1.81,3.61"
    assert not E.write_pattern_csv(q, dict(valid=False))
