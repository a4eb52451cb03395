"""Pins for the CPU oracle: each test checks an oracle function against something the oracle
does not compute itself — binary16 known answers, closed forms of the discrete Yee scheme,
conservation laws, and exact sums. A plausible mistake (wrong stagger, sign, index or operand)
fails at least one of them; tests/test_oracle_mutations.py compiles copies of the oracle with such
mistakes planted and shows that these pins reject each one."""
import ctypes as C
import math

import numpy as np
import pytest


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------------------------------------
# binary16 (src/SimulationGrid.cpp:256-264)
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("value,bits", [(6.3, 0x464D), (376.730313668 / 50.0, 0x4789),
                                        (2.0 / math.log(10.0), 0x3AF3), (6.0e4, 0x7B53), (1.0, 0x3C00),
                                        (0.0, 0x0000), (65504.0, 0x7BFF), (-2.5, 0xC100)])
def test_f16_known_answers(oracle, value, bits):
    assert oracle.lib().or_f32_to_f16(value) == bits


def test_f16_matches_numpy_rne(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(1)
    vals = np.concatenate([rng.standard_normal(4000) * 10.0 ** rng.integers(-9, 5, 4000),
                           (np.arange(0, 65536, dtype=np.uint16).view(np.float16)).astype(np.float32),
                           np.array([65519.0, 65520.0, 2.0 ** -25, 2.0 ** -24 * 1.5], np.float32)]).astype(np.float32)
    vals = vals[np.isfinite(vals) & (np.abs(vals) < 6.0e4)]
    ours = np.array([L.or_f32_to_f16(float(v)) for v in vals], np.uint16)
    ref = vals.astype(np.float16).view(np.uint16)
    assert np.array_equal(ours, ref)
    # decode round trip of every finite half
    halves = np.arange(0, 65536, dtype=np.uint16)
    fin = np.isfinite(halves.view(np.float16))
    back = np.array([L.or_f16_to_f32(int(h)) for h in halves[fin]], np.float32)
    assert np.array_equal(back, halves[fin].view(np.float16).astype(np.float32))


# ---------------------------------------------------------------------------------------------
# configure clamps (src/SimulationGrid.cpp:149-156)
# ---------------------------------------------------------------------------------------------
def test_configure_clamps(oracle):
    g = oracle.grid(4, 100, 30, 0.9, 80, -1.0, 50.0)
    assert (g.pml, g.nx, g.ny, g.nz) == (64, 136, 136, 136)
    assert g.dt == np.float32(np.float32(0.99) / np.float32(1.7320508))
    assert g.sigma_max == 0.0 and g.grading == 12.0
    g = oracle.grid(20, 20, 20, 0.01, -3, 1.6, 0.1)
    assert (g.pml, g.nx, g.dt, g.grading) == (0, 20, np.float32(0.05), np.float32(0.25))


# ---------------------------------------------------------------------------------------------
# absorber profile known values (shaders/common.glsl:23-43)
# ---------------------------------------------------------------------------------------------
def test_pml_profile_values(oracle):
    L = oracle.lib()
    g = oracle.grid(40, 40, 40, 0.5, 10, 1.6, 3.0)
    smax = 1.6
    # interior point: no absorber on any axis
    assert L.or_pml_sigma(C.byref(g), 20.0, 20.0, 20.0) == 0.0
    # the layer's inner edge (depth 0) and outer edge (depth L) on both sides
    assert L.or_pml_sigma(C.byref(g), 10.0, 20.0, 20.0) == 0.0
    assert L.or_pml_sigma(C.byref(g), 29.0, 20.0, 20.0) == 0.0
    assert L.or_pml_sigma(C.byref(g), 0.0, 20.0, 20.0) == pytest.approx(smax, rel=1e-6)
    assert L.or_pml_sigma(C.byref(g), 39.0, 20.0, 20.0) == pytest.approx(smax, rel=1e-6)
    # graded: depth 5 of 10 -> sigma_max (e^{g/2}-1)/(e^g-1); corners add the axes
    half = smax * (math.exp(1.5) - 1) / (math.exp(3.0) - 1)
    assert L.or_pml_sigma(C.byref(g), 5.0, 20.0, 20.0) == pytest.approx(half, rel=1e-6)
    assert L.or_pml_sigma(C.byref(g), 34.0, 5.0, 20.0) == pytest.approx(2 * half, rel=1e-6)
    # beyond the outer edge the profile saturates at sigma_max (u clamped to 1)
    assert L.or_pml_sigma(C.byref(g), 39.5, 20.0, 20.0) == pytest.approx(smax, rel=1e-6)
    # inPml: the +1/2 staggered samples of cell size-L-1 can be inside the layer
    assert L.or_in_pml(C.byref(g), 29, 20, 20) == 1 and L.or_in_pml(C.byref(g), 28, 20, 20) == 0
    assert L.or_in_pml(C.byref(g), 20, 9, 20) == 1 and L.or_in_pml(C.byref(g), 20, 10, 20) == 0


def test_phi_closed_form(oracle):
    L = oracle.lib()
    for x in [0.0, 1e-6, 1e-3, 9.999e-3, 1e-2, 0.3, 4.3, 3.4e4]:
        exact = 1.0 if x == 0 else -math.expm1(-x) / x
        # the exact branch (1 - e^{-x}) / x cancels in float: error ~ ulp(1) / x
        tol = 3e-7 + (1.2e-7 / x if x >= 1e-2 else 0.0)
        assert L.or_phi(x) == pytest.approx(exact, rel=tol, abs=0)
    # continuity across the branch switch at 1e-2
    lo, hi = L.or_phi(np.nextafter(np.float32(1e-2), np.float32(0))), L.or_phi(1e-2)
    assert abs(lo - hi) < 1e-6


# ---------------------------------------------------------------------------------------------
# Yee interior: exact discrete plane wave and the leapfrog energy invariant
# ---------------------------------------------------------------------------------------------
STAG_E = np.array([[0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5]])
STAG_B = np.array([[0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])


def plane_wave(size, k, E0, dt, n_e, n_b):
    """Discrete Yee plane wave: E at t = n_e dt, B at t = n_b dt, on EM_Lab's stagger
    (shaders/common.glsl:8-11). K = 2 sin(k/2), Omega = |K| = 2 sin(w dt/2)/dt, B0 = K x E0 / Omega."""
    nx, ny, nz = size
    K = 2.0 * np.sin(np.asarray(k) / 2.0)
    Om = np.linalg.norm(K)
    w = 2.0 / dt * math.asin(Om * dt / 2.0)
    B0 = np.cross(K, E0) / Om
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pos = np.stack([x, y, z], -1).reshape(-1, 3).astype(np.float64)
    E = np.zeros((pos.shape[0], 4))
    B = np.zeros((pos.shape[0], 4))
    for c in range(3):
        E[:, c] = np.real(E0[c] * np.exp(1j * ((pos + STAG_E[c]) @ k - w * n_e * dt)))
        B[:, c] = np.real(B0[c] * np.exp(1j * ((pos + STAG_B[c]) @ k - w * n_b * dt)))
    return E, B


def test_plane_wave_exact(oracle):
    size = (24, 22, 20)
    sim = oracle.OracleSim(size, 0.5, 0, sigma_max=0.0)
    n = np.prod(size)
    sim.upload_materials(np.zeros(n, np.float32), np.ones(n, np.float32))
    k = np.array([0.31, -0.23, 0.17])
    K = 2 * np.sin(k / 2)
    E0 = np.cross(K, [0.3, 0.5, -0.8])
    E0 = E0 / np.linalg.norm(E0)
    dt = float(sim.g.dt)
    E, B = plane_wave(size, k, E0, dt, 0, -0.5)
    sim.E[:, :3] = E[:, :3]
    sim.B[:, :3] = B[:, :3]
    steps = 6
    sim.step(steps)
    Ee, Be = plane_wave(size, k, E0, dt, steps, steps - 0.5)
    m = steps + 2
    idx = np.arange(n).reshape(size[2], size[1], size[0])[m:-m, m:-m, m:-m].ravel()
    assert np.max(np.abs(sim.E[idx, :3] - Ee[idx, :3])) < 2e-6
    assert np.max(np.abs(sim.B[idx, :3] - Be[idx, :3])) < 2e-6


def test_energy_invariant_cavity(oracle):
    """Lossless closed cavity (pml 0, sigma_max 0): |E^n|^2 + <B^{n+1/2}, B^{n-1/2}> is conserved
    only if update_E's curl is exactly the transpose of update_B's, boundaries included."""
    size = (11, 9, 10)
    sim = oracle.OracleSim(size, 0.55, 0, sigma_max=0.0)
    n = np.prod(size)
    sim.upload_materials(np.zeros(n, np.float32), np.ones(n, np.float32))
    rng = np.random.default_rng(3)
    sim.E[:, :3] = rng.standard_normal((n, 3)).astype(np.float32)
    sim.B[:, :3] = rng.standard_normal((n, 3)).astype(np.float32)
    inv = []
    for _ in range(60):
        e0 = sim.E[:, :3].astype(np.float64).copy()
        b0 = sim.B[:, :3].astype(np.float64).copy()
        sim.step(1)
        inv.append(np.sum(e0 * e0) + np.sum(sim.B[:, :3].astype(np.float64) * b0))
    inv = np.array(inv)
    assert np.max(np.abs(inv - inv[0])) / abs(inv[0]) < 2e-5


def test_lossy_medium_closed_form(oracle):
    """Uniform sigma, eps_r: a uniform E decays as exp(-sigma dt / eps) per step away from the
    walls; a uniform curl of B gives E^1 = (1 - exp(-x)) / sigma * curl."""
    size = (16, 16, 16)
    n = np.prod(size)
    L = oracle.lib()
    for sigma, eps in [(0.05, 2.0), (0.0004, 1.0), (3.0, 4.0)]:
        sim = oracle.OracleSim(size, 0.5, 0, sigma_max=0.0)
        sim.upload_materials(np.full(n, sigma, np.float32), np.full(n, eps, np.float32))
        sh = L.or_f16_to_f32(L.or_f32_to_f16(sigma))
        eh = L.or_f16_to_f32(L.or_f32_to_f16(eps))
        x = sh * 0.5 / eh
        sim.E[:, :3] = [1.0, -2.0, 0.5]
        sim.step(3)
        c = np.arange(n).reshape(16, 16, 16)[8, 8, 8]
        np.testing.assert_allclose(sim.E[c, :3], np.array([1.0, -2.0, 0.5]) * math.exp(-x) ** 3, rtol=2e-6)
        # Cb: Bz = g * y gives curl_x B = g everywhere except y = 0
        sim2 = oracle.OracleSim(size, 0.5, 0, sigma_max=0.0)
        sim2.upload_materials(np.full(n, sigma, np.float32), np.full(n, eps, np.float32))
        yy = (np.arange(n) // 16) % 16
        sim2.B[:, 2] = 0.25 * yy
        sim2.step(1)
        cb = -math.expm1(-x) / sh
        np.testing.assert_allclose(sim2.E[c, 0], 0.25 * cb, rtol=3e-6)


# ---------------------------------------------------------------------------------------------
# sources (shaders/apply_sources.comp, src/SimulationGrid.cpp:297-357)
# ---------------------------------------------------------------------------------------------
def _gauss_sum(c, s, offset):
    n = np.arange(-200, 400)
    return np.sum(np.exp(-(((n + offset) - c) / s) ** 2))


def test_point_source_footprint_sum(oracle):
    """E -= Cb J with Cb = dt in vacuum: the sum of Ez over the grid is -dt J0 * prod of 1-D
    Gaussian sums sampled at the Ez stagger (x, y integer; z + 1/2)."""
    from em_lab_amd.scene import Scene
    size = (40, 40, 40)
    sc = Scene(size=size, pml=0, courant=0.5)
    s = sc.hertzian((20.0, 20.3, 19.8), (0, 0, 1), 1.0, wavelength=16.0, waveform="cw", ramp=0.0)
    sim = oracle.OracleSim(size, 0.5, 0, sigma_max=0.0)
    n = np.prod(size)
    sim.upload_materials(np.zeros(n, np.float32), np.ones(n, np.float32))
    sim.set_sources([s])
    # one step: the signal at t = dt/2 is sin(w dt/2) (CW with no ramp)
    sim.step(1)
    sig = np.float32(math.sin(2 * math.pi / 16.0 * 0.25))
    J0 = np.float32(1.0) / (np.float32(math.pi) * np.float32(0.75) * np.float32(0.75))
    expect = -0.5 * float(J0) * float(sig) * _gauss_sum(20.0, 0.75, 0) * _gauss_sum(20.3, 0.75, 0) * \
        _gauss_sum(19.8, 0.75, 0.5)
    assert np.sum(sim.E[:, 2].astype(np.float64)) == pytest.approx(expect, rel=2e-5)
    # a z-polarised source leaves Ex, Ey untouched; only cells of the list are touched
    assert np.all(sim.E[:, :2] == 0)


def test_source_cell_list_covers_footprint(oracle):
    """Every cell where some staggered sample of a source footprint exceeds 1e-5 is listed; the
    list is sorted, duplicate-free and its entries index the right sources."""
    L = oracle.lib()
    from em_lab_amd.scene import Scene
    size = (30, 26, 28)
    sc = Scene(size=size, pml=0)
    srcs = [sc.hertzian((10.2, 12.7, 14.0), (0, 0, 1), 1.0), sc.hertzian((12.0, 12.0, 13.5), (1, 0, 0), 2.0)]
    g = oracle.grid(*size, 0.5, 0)
    arr = (oracle.OrSource * 2)(*[oracle.source_c(s) for s in srcs])
    nc, nl = C.c_int(), C.c_int()
    L.or_set_sources(C.byref(g), C.cast(arr, C.c_void_p), 2, None, None, None, C.byref(nc), None, C.byref(nl))
    cells = np.zeros((nc.value + 1, 4), np.int32)
    lists = np.zeros(nl.value, np.int32)
    packed = np.zeros((2, 16), np.float32)
    L.or_set_sources(C.byref(g), C.cast(arr, C.c_void_p), 2, _ptr(packed), None, _ptr(cells), C.byref(nc),
                     _ptr(lists), C.byref(nl))
    listed = {}
    for i in range(nc.value):
        key = tuple(cells[i, :3])
        listed[key] = set(lists[cells[i, 3]:cells[i + 1, 3]])
    lin = [(z * size[1] + y) * size[0] + x for (x, y, z) in listed]
    assert lin == sorted(lin) and len(set(lin)) == len(lin)
    for si, s in enumerate(srcs):
        c = np.array(s.center)
        for x in range(size[0]):
            for y in range(size[1]):
                for z in range(size[2]):
                    for st in STAG_E:
                        d = np.maximum(np.abs(np.array([x, y, z]) + st - c) - np.array(s.half_extent), 0) / 0.75
                        if math.exp(-np.dot(d, d)) > 1e-5:
                            assert si in listed.get((x, y, z), set())


def test_signal_closed_forms(oracle):
    L = oracle.lib()
    from em_lab_amd.emfdtd import SourceDesc
    pulse = SourceDesc(wavelength=20.0, waveform=1, ramp_periods=1.0, phase=0.0)
    tau = 20.0
    t0 = 3 * tau
    pc = oracle.source_c(pulse)
    # envelope peak at t0 (phase 0 -> the sine is 0 there); quarter period later: exp(-(5/20)^2)
    assert L.or_signal(C.byref(pc), t0) == 0.0
    assert L.or_signal(C.byref(pc), t0 + 5.0) == pytest.approx(math.exp(-(5.0 / tau) ** 2), rel=1e-13)
    assert L.or_signal(C.byref(pc), t0 + 6.01 * tau) == 0.0  # |u| > 6 cut
    assert L.or_signal(C.byref(pc), -1.0) == 0.0
    cw = SourceDesc(wavelength=16.0, waveform=0, ramp_periods=3.0, phase=0.7)
    cc = oracle.source_c(cw)
    T = 48.0
    for t in [0.0, 7.0, 24.0, 47.9, 48.0, 100.0]:
        r = min(t / T, 1.0)
        ramp = r ** 3 * (10 - 15 * r + 6 * r * r)
        assert L.or_signal(C.byref(cc), t) == pytest.approx(ramp * math.sin(2 * math.pi / 16 * t + float(np.float32(0.7))),
                                                            abs=1e-13)
    rep = SourceDesc(wavelength=10.0, waveform=1, ramp_periods=0.5, repeat_periods=7.0, delay_periods=1.0)
    rc = oracle.source_c(rep)
    # repetition interval max(7 periods, 2 t0) = 70; delay 1 period
    assert L.or_signal(C.byref(rc), 10.0 + 3.3) == pytest.approx(L.or_signal(C.byref(rc), 10.0 + 73.3), abs=1e-12)


# ---------------------------------------------------------------------------------------------
# monitors
# ---------------------------------------------------------------------------------------------
def test_two_stage_ema_converges_to_phasor(oracle):
    """Port V(t) = A cos(w t + p): stage 2 tends to A e^{ip} with ripple ~ 1/(w tau)^2."""
    L = oracle.lib()
    size = (8, 8, 8)
    g = oracle.grid(*size, 0.5, 0)
    n = 512
    E = np.zeros((n, 4), np.float32)
    B = np.zeros((n, 4), np.float32)
    port = (oracle.OrPort * 1)(oracle.OrPort((C.c_int * 3)(3, 3, 3), 2, 1, 1))
    ph = np.zeros(8, np.float32)
    A, p, wl = 0.8, 0.6, 16.0
    w = 2 * math.pi / wl
    idx = (3 * 8 + 3) * 8 + 3
    rot = np.zeros(4, np.float32)
    blend = C.c_float()
    for step in range(1, 1600):
        t = step * 0.5
        E[idx, 2] = A * math.cos(w * t + p)
        # B chosen so that the loop current is I(t - dt/2) = 2 A cos(w (t - dt/2) - p)
        B[idx, 1] = 2 * A * math.cos(w * (t - 0.25) - p)
        L.or_monitor_constants(wl, 4.0, 0.5, step, _ptr(rot), C.byref(blend))
        L.or_port_accumulate(C.byref(g), _ptr(E), _ptr(B), C.cast(port, C.c_void_p), 1, _ptr(rot), blend.value, -1,
                             _ptr(ph), None)
    assert blend.value == pytest.approx(1 - math.exp(-0.5 / 32.0), rel=1e-6)
    assert complex(ph[4], ph[5]) == pytest.approx(A * np.exp(1j * p), abs=1.5e-3)
    assert complex(ph[6], ph[7]) == pytest.approx(2 * A * np.exp(-1j * p), abs=3e-3)


def test_sweep_resistor_closed_form(oracle):
    """With I_k = -V_k / R the sweep gives Z = R eta0 exp(-i w dt / 2) exactly (the half-step
    factor applied to I), S11 from it, and the recurrence matches a direct DFT."""
    L = oracle.lib()
    dt = 0.5
    n = 3000
    t = np.arange(n) * dt
    V = (np.exp(-((t - 60) / 20) ** 2) * np.sin(2 * math.pi / 20 * (t - 60))).astype(np.float32)
    Rn = 50.0 / 376.730313668
    I = (-V / np.float32(Rn)).astype(np.float32)
    series = np.stack([V, I], -1).astype(np.float32).reshape(n, 1, 2)
    out = np.zeros((16, 5))
    band = np.zeros(5)
    L.or_sweep(_ptr(series), n, 1, 0, dt, 13.0, 40.0, 16, 50.0, _ptr(out), _ptr(band))
    nu = 1 / 40 + (1 / 13 - 1 / 40) * np.arange(16) / 15
    w = 2 * math.pi * nu
    Z = np.float32(Rn) * 376.730313668 * np.exp(-0.5j * w * dt)
    # exact up to the float32 rounding of I = -V / Rn
    np.testing.assert_allclose(out[:, 1] + 1j * out[:, 2], Z, rtol=2e-7)
    np.testing.assert_allclose(out[:, 0], 1 / nu, rtol=1e-14)
    g = np.abs((Z - 50) / (Z + 50))
    np.testing.assert_allclose(out[:, 3], 20 * np.log10(g), atol=1e-5)
    # direct DFT against the renormalised rotation recurrence
    Vd = np.sum(V.astype(np.float64)[None, :] * np.exp(-1j * w[:, None] * t[None, :]), 1)
    Id = np.sum(I.astype(np.float64)[None, :] * np.exp(-1j * w[:, None] * t[None, :]), 1) * np.exp(0.5j * w * dt)
    np.testing.assert_allclose(out[:, 1] + 1j * out[:, 2], -Vd / Id * 376.730313668, rtol=1e-9)


def test_flux_uniform_fields(oracle):
    """Uniform E = (0, e, 0) and B = (0, 0, b): S = (e b, 0, 0) at every node; a plane of
    U x V cells carries U V e b; a closed box carries 0 net."""
    size = (20, 18, 16)
    sim = oracle.OracleSim(size, 0.5, 0, sigma_max=0.0)
    sim.E[:, 1] = 0.7
    sim.B[:, 2] = 1.3
    assert sim.plane_flux(0, 10, (2, 3), (12, 9)) == pytest.approx(10 * 6 * 0.7 * 1.3, rel=1e-6)
    assert sim.plane_flux(1, 10, (2, 3), (12, 9)) == pytest.approx(0.0, abs=1e-9)
    assert sim.box_flux((3, 3, 3), (12, 11, 10)) == pytest.approx(0.0, abs=1e-6)
    # linear field Ey = x: the box sees (x_max - x_min) e b per node of the x faces
    x = np.arange(np.prod(size)) % size[0]
    sim.E[:, 1] = x.astype(np.float32)
    ny, nz = 11 - 3 + 1, 10 - 3 + 1
    assert sim.box_flux((3, 3, 3), (12, 11, 10)) == pytest.approx((12 - 3) * ny * nz * 1.3, rel=1e-6)


def test_probe_linear_fields_exact(oracle):
    """Trilinear interpolation reproduces a linear field exactly when each component is sampled
    at its own Yee stagger (src/FieldAnalyzer.cpp:300-306)."""
    size = (12, 11, 10)
    sim = oracle.OracleSim(size, 0.5, 0)
    n = np.prod(size)
    z, y, x = np.unravel_index(np.arange(n), (size[2], size[1], size[0]))
    pos = np.stack([x, y, z], -1).astype(np.float64)
    coef = np.array([[0.3, -0.2, 0.1], [0.05, 0.4, -0.3], [-0.1, 0.2, 0.25]])
    for c in range(3):
        sim.E[:, c] = (pos + STAG_E[c]) @ coef[c] + c
        sim.B[:, c] = (pos + STAG_B[c]) @ coef[(c + 1) % 3] - c
    for p in [(5.3, 4.7, 3.2), (6.0, 6.5, 5.25)]:
        out = sim.probe(p)
        for c in range(3):
            assert out[c] == pytest.approx(np.dot(p, coef[c]) + c, abs=2e-5)
            assert out[3 + c] == pytest.approx(np.dot(p, coef[(c + 1) % 3]) - c, abs=2e-5)
        np.testing.assert_allclose(out[6:9], np.cross(out[0:3], out[3:6]), rtol=1e-5, atol=1e-6)


def test_directions_known_structure(oracle):
    L = oracle.lib()
    dirs = np.zeros((10590, 4), np.float32)
    mi = np.zeros(2 * 81 * 81, np.int32)
    fr = np.zeros(9, np.float32)
    axis = np.array([0, 0, 1], np.float32)
    nd = L.or_build_directions(_ptr(axis), _ptr(dirs), dirs.shape[0], _ptr(mi), _ptr(fr))
    # 3 cuts x 180, plus both hemispheres: every (i, j) in [-40, 40]^2 with i^2 + j^2 <= 1600
    # (Gauss circle count 5025; all 12 rim points pass u*u + v*v <= 1 in IEEE float)
    gauss = sum(1 for i in range(-40, 41) for j in range(-40, 41) if i * i + j * j <= 1600)
    assert gauss == 5025 and nd == 540 + 2 * gauss
    np.testing.assert_allclose(np.linalg.norm(dirs[:, :3], axis=1), 1.0, atol=1e-6)
    np.testing.assert_allclose(dirs[0, :3], [1, 0, 0], atol=1e-7)
    np.testing.assert_allclose(dirs[45, :3], [0, 1, 0], atol=1e-6)      # xy cut, 90 deg
    np.testing.assert_allclose(dirs[180 + 45, :3], [0, 0, 1], atol=1e-6)  # xz cut, 90 deg
    np.testing.assert_allclose(dirs[360 + 45, :3], [0, 0, 1], atol=1e-6)  # yz cut, 90 deg
    centre = mi[40 * 81 + 40]
    np.testing.assert_allclose(dirs[centre, :3], [0, 0, 1], atol=1e-7)
    assert (mi[:81 * 81] >= 0).sum() == 5025
    # the lower hemisphere mirrors the upper one through the map plane
    lo = mi[81 * 81:][mi[81 * 81:] >= 0]
    up = mi[:81 * 81][mi[:81 * 81] >= 0]
    np.testing.assert_allclose(dirs[lo, :2], dirs[up, :2])
    np.testing.assert_allclose(dirs[lo, 2], -dirs[up, 2])


def _hertzian_phasors(bmin, bmax, fo, npts, k, center, oracle):
    """Exact near field of a z-directed Hertzian dipole (I l = 1, eta = 1) at every box node,
    written into stage-2 phasor slots. E ~ e^{-ikr}; the far field ~ sin(theta)."""
    L = oracle.lib()
    ph = np.zeros((npts, 32), np.float32)
    node = np.zeros(3, np.int32)
    nrm = np.zeros(3, np.float32)
    w = C.c_float()
    for i in range(npts):
        L.or_ntff_point(_ptr(bmin), _ptr(bmax), _ptr(fo), i, _ptr(node), _ptr(nrm), C.byref(w))
        r = node.astype(np.float64) - center
        R = np.linalg.norm(r)
        rh = r / R
        ct = rh[2]
        g = np.exp(-1j * k * R) / (4 * np.pi * R)
        # H = ik (1 + 1/(ikR)) g (z x r) ; E = from the dipole formulas (eta = 1)
        zxr = np.cross([0, 0, 1], rh)
        H = 1j * k * (1 + 1 / (1j * k * R)) * g * zxr
        Er = 2 * (1 / R + 1 / (1j * k * R * R)) * g * ct
        Et = 1j * k * (1 + 1 / (1j * k * R) - 1 / (k * k * R * R)) * g * np.sqrt(max(1 - ct * ct, 0))
        rho = np.linalg.norm(r[:2])
        th_hat = np.array([r[0] * r[2] / (R * rho), r[1] * r[2] / (R * rho), -rho / R]) if rho > 0 else np.zeros(3)
        E = Er * rh + Et * th_hat
        # these are e^{+i w t} phasors (outgoing e^{-ikR}), the monitor's convention: the EMA of
        # 2 F(t) e^{-i w t} recovers A for F(t) = Re{A e^{+i w t}} (shaders/ntff_accumulate.comp:5-7)
        ph[i, 16:19] = E.real
        ph[i, 20:23] = E.imag
        ph[i, 24:27] = H.real
        ph[i, 28:31] = H.imag
    return ph


def test_ntff_hertzian_directivity(oracle):
    """Exact Hertzian near field on a Huygens box -> D = 1.5 sin^2(theta) (1.76 dBi), HPBW 90."""
    L = oracle.lib()
    bmin = np.array([0, 0, 0], np.int32)
    bmax = np.array([24, 24, 24], np.int32)
    fo = np.zeros(7, np.int32)
    npts = L.or_ntff_face_offsets(_ptr(bmin), _ptr(bmax), _ptr(fo))
    assert npts == 6 * 25 * 25
    wl = 16.0
    k = 2 * np.pi / wl
    center = np.array([12.0, 12.0, 12.0])
    ph = _hertzian_phasors(bmin, bmax, fo, npts, k, center, oracle)
    dirs = np.zeros((10590, 4), np.float32)
    mi = np.zeros(2 * 81 * 81, np.int32)
    fr = np.zeros(9, np.float32)
    axis = np.array([0, 0, 1], np.float32)
    L.or_build_directions(_ptr(axis), _ptr(dirs), 10590, _ptr(mi), _ptr(fr))
    out = np.zeros(2 * 10590, np.float32)
    L.or_ntff_eval_f32(_ptr(bmin), _ptr(bmax), _ptr(fo), npts, _ptr(ph), _ptr(dirs), 10590, np.float32(k),
                       _ptr(center.astype(np.float32)), _ptr(out))
    P = float(out[1])
    assert P == pytest.approx(k * k / (12 * np.pi), rel=0.02)  # Hertzian P = eta k^2 (I l)^2 / (12 pi)
    D = k * k / (8 * np.pi * P) * out[0::2].astype(np.float64)
    sin2 = 1 - dirs[:, 2].astype(np.float64) ** 2
    np.testing.assert_allclose(D, 1.5 * sin2, atol=0.03)
    # f32 reference order against an f64 brute-force sum
    f2 = np.zeros(200)
    P64 = C.c_double()
    d64 = np.ascontiguousarray(dirs[::53, :3].astype(np.float64))
    L.or_ntff_eval_f64(_ptr(bmin), _ptr(bmax), _ptr(fo), npts, _ptr(ph), _ptr(d64), d64.shape[0], k,
                       _ptr(center), _ptr(f2), C.byref(P64))
    np.testing.assert_allclose(out[0::2][::53], f2[:d64.shape[0]], rtol=2e-4, atol=1e-7 * f2.max())
    assert P == pytest.approx(P64.value, rel=1e-5)


# ---------------------------------------------------------------------------------------------
# edge averaging of uploadMaterials (src/SimulationGrid.cpp:224-248)
# ---------------------------------------------------------------------------------------------
def _unpack_edges(oracle, edge):
    """(n, 3) edge sigma and (n, 3) edge eps_r from the packed half2 edge texels."""
    L = oracle.lib()
    lo = (edge & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float64)
    hi = (edge >> 16).astype(np.uint16).view(np.float16).astype(np.float64)
    sig = np.stack([lo[:, 0], hi[:, 0], lo[:, 1]], 1)
    eps = np.stack([hi[:, 1], lo[:, 2], hi[:, 2]], 1)
    assert L.or_f16_to_f32(L.or_f32_to_f16(1.0)) == 1.0
    return sig, eps


# Yee edge positions relative to the texel (i, j, k): Ex (i+1/2, j, k), Ey (i, j+1/2, k), Ez (i, j, k+1/2)
EDGE_AT = np.array([[0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5]])


def test_edge_average_linear_medium(oracle):
    """sigma and eps_r linear in (x, y, z), sampled at the cell centres (i+1/2, j+1/2, k+1/2): the
    4-cell average on a Yee edge is the field at the edge midpoint, exactly (values and sums are
    binary16- and float-exact). At the lower faces the neighbour index is clamped to 0, so the
    transverse coordinate reads 1/2 instead of 0 there."""
    size = (9, 8, 10)  # the grid is at least 8 texels wide (configure clamp)
    nx, ny, nz = size
    n = nx * ny * nz
    z, y, x = np.unravel_index(np.arange(n), (nz, ny, nx))
    pos = np.stack([x, y, z], -1).astype(np.float64)
    cs, ce = np.array([0.25, 0.125, 0.5]), np.array([0.5, 0.25, 0.125])
    sig = (0.5 + (pos + 0.5) @ cs).astype(np.float32)
    eps = (1.0 + (pos + 0.5) @ ce).astype(np.float32)
    sim = oracle.OracleSim(size, 0.5, 0, sigma_max=0.0)
    sim.upload_materials(sig, eps)
    es, ee = _unpack_edges(oracle, sim.edge)
    for c in range(3):
        at = pos + EDGE_AT[c]
        # transverse axes: clamped neighbour at index 0 -> both cells sit at coordinate 1/2
        for a in range(3):
            if a != c:
                at[:, a] = np.maximum(at[:, a], 0.5)
        np.testing.assert_array_equal(es[:, c], 0.5 + at @ cs, err_msg=f"sigma, component {c}")
        np.testing.assert_array_equal(ee[:, c], 1.0 + at @ ce, err_msg=f"eps_r, component {c}")


@pytest.mark.parametrize("cell", [(4, 3, 5), (0, 3, 2), (3, 0, 0), (0, 0, 0)])
def test_edge_average_isolated_cell(oracle, cell):
    """One non-vacuum cell touches exactly the 4 Ex, 4 Ey and 4 Ez edges its cube has (a quarter
    of its value each); an edge on a lower face at index 0 counts the cell once more for each
    clamped neighbour (1/2 or all of it)."""
    size = (8, 9, 10)
    nx, ny, nz = size
    n = nx * ny * nz
    a, b, c0 = cell
    sig = np.zeros(n, np.float32)
    eps = np.ones(n, np.float32)
    idx = (c0 * ny + b) * nx + a
    sig[idx] = 8.0
    eps[idx] = 5.0
    sim = oracle.OracleSim(size, 0.5, 0, sigma_max=0.0)
    sim.upload_materials(sig, eps)
    es, ee = _unpack_edges(oracle, sim.edge)
    es = es.reshape(nz, ny, nx, 3)
    ee = ee.reshape(nz, ny, nx, 3)
    for comp in range(3):
        want = np.zeros((nz, ny, nx))
        # the edge of component comp at texel t averages cells t - {0,1} on the two other axes
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    d = (dx, dy, dz)
                    if d[comp] != 0:
                        continue
                    # cell (a, b, c0) is the neighbour t - d of texel t = cell + d (clamped at 0)
                    for tz in range(nz):
                        for ty in range(ny):
                            for tx in range(nx):
                                t = (tx, ty, tz)
                                nb = tuple(max(t[i] - d[i], 0) for i in range(3))
                                if nb == (a, b, c0):
                                    want[tz, ty, tx] += 0.25
        assert np.count_nonzero(want) > 0
        np.testing.assert_array_equal(es[..., comp], 8.0 * want, err_msg=f"sigma comp {comp}")
        np.testing.assert_array_equal(ee[..., comp], 1.0 + 4.0 * want, err_msg=f"eps comp {comp}")
        if cell == (4, 3, 5):
            # interior cell: exactly four edges per component, each holding a quarter
            assert np.count_nonzero(want) == 4 and np.all(want[want > 0] == 0.25)
            touched = set(zip(*np.nonzero(want)))
            axes = [i for i in range(3) if i != comp]
            expect = set()
            for u in (0, 1):
                for v in (0, 1):
                    t = [a, b, c0]
                    t[axes[0]] += u
                    t[axes[1]] += v
                    expect.add((t[2], t[1], t[0]))
            assert touched == expect


# ---------------------------------------------------------------------------------------------
# Huygens-node sampling eAtNode / bAtNode (shaders/field_sampling.glsl:17-32)
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("seed", [0, 1])
def test_node_sampling_linear_fields(oracle, seed):
    """A field linear in position, stored at each component's Yee stagger, is returned at the
    integer node exactly: the E means run along the component's own axis, the B means over the two
    transverse axes. At index 0 the clamped neighbour is the texel itself, so the node value is
    the staggered sample's own (coordinate 1/2 on the averaged axes)."""
    L = oracle.lib()
    size = (10, 9, 8)
    nx, ny, nz = size
    n = nx * ny * nz
    rng = np.random.default_rng(seed)
    ce = rng.integers(-8, 9, (3, 3)) / 16.0
    cb = rng.integers(-8, 9, (3, 3)) / 16.0
    z, y, x = np.unravel_index(np.arange(n), (nz, ny, nx))
    pos = np.stack([x, y, z], -1).astype(np.float64)
    sim = oracle.OracleSim(size, 0.5, 0)
    for c in range(3):
        sim.E[:, c] = (pos + STAG_E[c]) @ ce[c] + c
        sim.B[:, c] = (pos + STAG_B[c]) @ cb[c] - c
    out = np.zeros(3, np.float32)
    for p in [(4, 5, 3), (9, 8, 7), (1, 1, 1), (0, 4, 2), (3, 0, 0), (0, 0, 0)]:
        L.or_e_at_node(C.byref(sim.g), _ptr(sim.E), *p, _ptr(out))
        e = out.copy()
        L.or_b_at_node(C.byref(sim.g), _ptr(sim.B), *p, _ptr(out))
        b = out.copy()
        q = np.array(p, np.float64)
        for c in range(3):
            qe = q.copy()
            if q[c] == 0:
                qe[c] = 0.5
            assert e[c] == np.dot(qe, ce[c]) + c, (p, c, "E")
            qb = q.copy()
            for a in range(3):
                if a != c and q[a] == 0:
                    qb[a] = 0.5
            assert b[c] == np.dot(qb, cb[c]) - c, (p, c, "B")
