"""The opt-in convolutional PML on the GPU (emfd_grid_set_absorber(EMFD_ABSORBER_CPML)):

- the same fields as the oracle's restatement, bit for bit, with lossy materials reaching into
  the slabs and a source inside them, on the vectorised and the scalar two-pass kernels and on
  the one-pass march with its slab corrections; the march also with thin-wire dipoles and feeds
  in and next to the slabs, ports and a far-field box bordering the slabs, against the two passes;
- the closed form of a stretched-coordinate layer (cpml_reflection.py) at 16 cells: the
  normal-incidence level, the cos(theta) law of oblique incidence (ln R(theta) = cos(theta)
  ln R(0), which a graded conductivity does not obey), and the low-frequency shift of the
  complex-frequency-shifted stretch (ln R scales with the integral of sigma w^2 / (alpha^2 + w^2))."""
import numpy as np
import pytest

from em_lab_amd import emfdtd as E
from em_lab_amd.scene import MAX_COURANT, Scene

from cpml_reflection import closed_ln_r, measured_r, record_port, sheet_sources
from gpu_helpers import gpu_grid, oracle_sim, rel_l2

pytestmark = pytest.mark.gpu

L, GRADING = 16, 3.0
SIGMA_MAX = 0.3  # the layer's integral 1.5: R = e^-3 at normal incidence
LAM, PW = 10.0, 0.4
LAMS = [10.0, 12.0, 16.0, 20.0]


def _lossy_scene(size, pml, seed):
    sc = Scene(size=size, pml=pml, courant=0.55)
    rng = np.random.default_rng(seed)
    mask = rng.random(sc.sigma.shape) < 0.2
    sc.sigma[mask] = rng.uniform(0.0, 1.0, int(mask.sum())).astype(np.float32)
    sc.eps[mask] = rng.choice(np.array([2.0, 4.0], np.float32), int(mask.sum()))
    c = [v / 2 for v in size]
    sc.hertzian((c[0] + 0.3, c[1] - 0.2, c[2] + 0.1), (0.4, -0.3, 1.0), 1.0, wavelength=9.0, waveform="pulse",
                pulse_width=0.6)
    sc.hertzian((pml + 1.5, c[1], c[2]), (1, 0, 0), 0.5, wavelength=7.0, waveform="cw", ramp=1.0)
    return sc


@pytest.mark.parametrize("size,kernel", [((40, 36, 32), E.KERNEL_AUTO), ((37, 33, 30), E.KERNEL_AUTO),
                                         ((40, 36, 32), E.KERNEL_TWO_PASS), ((40, 36, 32), E.KERNEL_MARCH),
                                         ((44, 40, 33), E.KERNEL_MARCH)])
def test_cpml_matches_oracle(size, kernel):
    print("case: cpml vs oracle", size, kernel, flush=True)
    sc = _lossy_scene(size, 6, sum(size))
    g = E.SimulationGrid(0)
    g.configure(sc.grid_config())
    s, e = sc.flat_materials()
    g.upload_materials(s, e, sc.subcell)
    g.set_sources(sc.sources)
    g.set_kernel(kernel)
    g.set_absorber(E.ABSORBER_CPML, 1.2, 3.0, 0.15)
    sim = oracle_sim(sc, ports=False, ntff=False)
    sim.set_cpml(1.2, 3.0, 0.15)
    g.step(90)
    sim.step(90)
    Eg, Bg = g.fields(E.LAYOUT_TEXEL)
    re, rb = rel_l2(Eg[:, :3], sim.E[:, :3]), rel_l2(Bg[:, :3], sim.B[:, :3])
    print(f"rel L2 E {re:.2e} B {rb:.2e}, bit-identical {np.array_equal(Eg[:, :3], sim.E[:, :3])}", flush=True)
    assert np.abs(Eg[:, :3]).max() > 0
    assert np.array_equal(Eg[:, :3], sim.E[:, :3]) and np.array_equal(Bg[:, :3], sim.B[:, :3])
    g.close()


def _wire_scene():
    sc = Scene(size=(72, 64, 56), pml=8, courant=0.55)
    rng = np.random.default_rng(5)
    mask = rng.random(sc.sigma.shape) < 0.03
    sc.sigma[mask] = rng.uniform(0.0, 0.5, int(mask.sum())).astype(np.float32)
    sc.eps[mask] = 3.0
    sc.box((2, 20, 20), (12, 30, 34), eps_r=2.0, sigma=0.05)  # across the low x slab's inner face
    # a dipole inside the low x slab, one across the high z slab's face, one inside the box
    sc.dipole_thin((5, 32, 28), axis=2, length=7.0, radius=0.1, impedance=50.0, amplitude=1.0, wavelength=14.0,
                   waveform="pulse", pulse_width=1.0)
    sc.dipole_thin((36, 32, 46), axis=2, length=9.0, radius=0.1, impedance=50.0, amplitude=0.5, wavelength=16.0,
                   waveform="pulse", pulse_width=1.0)
    sc.dipole_thin((36, 30, 26), axis=1, length=9.0, radius=0.1, impedance=75.0, amplitude=1.0, wavelength=16.0,
                   waveform="pulse", pulse_width=1.0)
    sc.hertzian((66.5, 9.2, 50.1), (0.3, 1.0, 0.2), 0.7, wavelength=9.0, waveform="pulse", pulse_width=0.7)
    sc.emit_thin_wires()
    sc.set_far_field(margin=0, wavelength=16.0)
    return sc


def _run_cpml(sc, kernel, steps):
    g, pm, ff = gpu_grid(sc, kernel)
    g.set_absorber(E.ABSORBER_CPML, 1.2, 3.0, 0.15)
    g.step(steps)
    Eg, Bg = g.fields(E.LAYOUT_SOA)
    out = {"E": Eg, "B": Bg, "ports": pm.phasors(), "ntff": ff.phasors()}
    g.close()
    return out


def test_cpml_march_wires_ports_far_field():
    """Thin-wire dipoles and resistive feeds inside and across the slabs, lossy material texels in
    them, a Hertzian source in a slab corner and a far-field box whose faces border the slabs: the
    march with its slab corrections gives the fields, port and far-field phasors of the two passes
    with the per-axis corrections, bit for bit."""
    print("case: cpml march vs two passes, wires in the slabs", flush=True)
    sc = _wire_scene()
    a = _run_cpml(sc, E.KERNEL_MARCH, 150)
    b = _run_cpml(sc, E.KERNEL_TWO_PASS4, 150)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert np.abs(a["E"]).max() > 0 and np.abs(a["ntff"]).max() > 0


def _layout(theta_deg):
    xs = L + 22
    xo = xs + 6
    nx = xo + 10 + L + 1
    return xs, xo, nx


def _record(nx, ny, nz, xs, xo, theta, pol, steps, alpha_max=0.0):
    g = E.SimulationGrid(0)
    g.configure(E.GridConfig((nx, ny, nz), MAX_COURANT, L))
    g.set_sources(sheet_sources(ny, nz, xs, LAM, PW, theta, pol))
    g.set_absorber(E.ABSORBER_CPML, SIGMA_MAX, GRADING, alpha_max)
    pm = E.PortMonitor(g)
    pm.configure([record_port((xo, ny // 2, nz // 2), pol, LAM)], 4.0, None, 1e-3, record_steps=steps)
    g.step(steps)
    v = pm.read_series()[:, 0, 0].astype(np.float64)
    g.close()
    return v


def _reflection(theta, alpha_max=0.0):
    xs, xo, nx = _layout(theta)
    ny, nz = (250, 160) if theta else (160, 160)
    pol = 2 if theta else 1
    c, s = np.cos(np.radians(theta)), np.sin(np.radians(theta))
    t_inc = 3 * PW * LAM + (xo - xs) * c + (ny / 2) * s  # incident peak at the observation edge
    t_ref = t_inc + 2 * (nx - xo) * c                     # reflected peak
    steps = int((t_ref + 3.5 * PW * LAM) / MAX_COURANT) + 10
    vt = _record(nx, ny, nz, xs, xo, theta, pol, steps, alpha_max)
    vr = _record(nx + steps, ny, nz, xs, xo, theta, pol, steps, alpha_max)
    return measured_r(vt, vr, MAX_COURANT, 0.5 * (t_inc + t_ref), LAMS)


def test_cpml_normal_incidence_closed_form():
    print("case: cpml normal incidence, 16-cell layer", flush=True)
    r = _reflection(0.0)
    ratio = np.array([np.log(v) / closed_ln_r(SIGMA_MAX, GRADING, L, lam) for v, lam in zip(r, LAMS)])
    print("R", r, "ln R / closed form", ratio, flush=True)
    assert np.all(ratio > 0.95) and np.all(ratio < 1.15)


@pytest.mark.parametrize("theta", [20.0, 30.0])
def test_cpml_oblique_cos_law(theta):
    print("case: cpml oblique incidence", theta, flush=True)
    r0 = _reflection(0.0)
    rt = _reflection(theta)
    law = np.log(rt) / np.log(r0)
    print("R(0)", r0, "R(theta)", rt, "ln ratio", law, "cos", np.cos(np.radians(theta)), flush=True)
    assert np.all(np.abs(law / np.cos(np.radians(theta)) - 1.0) <= 0.06)


def test_cpml_frequency_shift():
    print("case: cpml complex-frequency shift alpha_max 0.3", flush=True)
    r0 = _reflection(0.0)
    ra = _reflection(0.0, alpha_max=0.3)
    got = np.log(ra) / np.log(r0)
    want = np.array([closed_ln_r(SIGMA_MAX, GRADING, L, lam, alpha_max=0.3) / closed_ln_r(SIGMA_MAX, GRADING, L, lam)
                     for lam in LAMS])
    print("ln R(alpha) / ln R(0):", got, "closed form", want, flush=True)
    assert np.all(np.abs(got - want) <= 0.05)
    assert got[-1] < got[0]  # the shift spares low frequencies most
