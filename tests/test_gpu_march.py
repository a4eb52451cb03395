"""The one-pass march kernel (EMFD_KERNEL_MARCH): bit-identical fields, port and far-field
phasors to the two-pass kernel (itself bit-identical to the oracle), and the oracle directly on
the plain Hertzian scene. Covers several x segments (nx > 256) with lane-0 / lane-63 halos at the
segment seams, z-chunk seams, row counts no group divides, materials inside the absorber and thin
wires next to the seams (the sparse recomputation of their texels)."""
import numpy as np
import pytest

from em_lab_amd import emfdtd as E
from em_lab_amd.scene import Scene, hertzian_vacuum, spheres_dipole

from gpu_helpers import absorber_material_scene, gpu_grid, oracle_sim, rel_l2

pytestmark = pytest.mark.gpu


def _run(scene, kernel, steps, ports=True, ntff=True):
    g, pm, ff = gpu_grid(scene, kernel, ports=ports, ntff=ntff)
    g.step(steps)
    Eg, Bg = g.fields(E.LAYOUT_SOA)
    out = {"E": Eg, "B": Bg}
    if pm is not None:
        out["ports"] = pm.phasors()
    if ff is not None:
        out["ntff"] = ff.phasors()
    g.close()
    return out


def _same(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_march_hertzian_64_oracle():
    """64^3 vacuum, 8-cell absorber, z Hertzian pulse, 400 steps: fields and far-field phasors
    against the oracle (the same bound as every kernel) and bit-identical to the two-pass."""
    print("case: march hertzian 64^3", flush=True)
    sc = hertzian_vacuum(64, 8)
    g, _, ff = gpu_grid(sc, E.KERNEL_MARCH)
    sim = oracle_sim(sc, gpu_ntff_setup=ff.setup)
    g.step(400)
    sim.step(400)
    Eg, Bg = g.fields(E.LAYOUT_TEXEL)
    assert rel_l2(Eg[:, :3], sim.E[:, :3]) <= 1e-5 and rel_l2(Bg[:, :3], sim.B[:, :3]) <= 1e-5
    assert rel_l2(ff.phasors(), sim.ntff["phasors"]) <= 1e-5
    assert np.abs(Eg[:, :3]).max() > 0
    g.close()
    _same(_run(sc, E.KERNEL_MARCH, 150), _run(sc, E.KERNEL_AUTO, 150))


def test_march_spheres_thin_dipole_bit_identical():
    """Dielectric spheres and a thin-wire dipole behind a resistive feed: every material, feed
    and wire texel is recomputed by the sparse passes; fields, port and far-field phasors equal
    the two-pass kernel's bit for bit."""
    print("case: march spheres + dipole 48^3", flush=True)
    sc = spheres_dipole(n=48, pml=6, n_spheres=5, wavelength=20.0, radius_range=(4.0, 10.0), sweep_points=16)
    _same(_run(sc, E.KERNEL_MARCH, 200), _run(sc, E.KERNEL_AUTO, 200))


@pytest.mark.parametrize("zchunk", [0, 3, 7])
def test_march_segments_and_seams(zchunk):
    """nx = 520 (three x segments, the last 8 cells wide), ny = 23 (no multiple of the group's 7
    rows), short z-chunks; random lossy materials, a box reaching into the absorber, a thin wire
    and a feed whose texels sit at the x = 255 / 256 segment seam, two sources."""
    print("case: march seams zchunk", zchunk, flush=True)
    size = (520, 23, 20)
    sc = Scene(size=size, pml=3, courant=0.55)
    rng = np.random.default_rng(11 + zchunk)
    mask = rng.random(sc.sigma.shape) < 0.05
    sc.sigma[mask] = rng.uniform(0.0, 1.5, int(mask.sum())).astype(np.float32)
    sc.eps[mask] = rng.choice(np.array([2.0, 5.0], np.float32), int(mask.sum()))
    sc.box((1, 5, 5), (8, 12, 15), eps_r=3.0, sigma=0.2)
    sc.hertzian((255.5, 11.3, 9.7), (0.2, 0.6, 1.0), 1.0, wavelength=12.0, waveform="pulse", pulse_width=0.8)
    sc.hertzian((512.0, 4.0, 15.0), (1, 0, 0), 0.5, wavelength=9.0, waveform="cw", ramp=1.0)
    sc.dipole_thin((256, 11, 10), axis=2, length=7.0, radius=0.1, impedance=50.0, amplitude=1.0, wavelength=14.0,
                   waveform="pulse", pulse_width=1.0)
    sc.emit_thin_wires()
    a = _run(sc, E.KERNEL_MARCH | (zchunk << 8), 60, ntff=False)
    b = _run(sc, E.KERNEL_AUTO, 60, ntff=False)
    _same(a, b)
    assert np.abs(a["E"]).max() > 0


def test_march_material_in_absorber():
    """Dielectric boxes reaching into the absorber (class texels with the absorber's
    conductivity added: the recomputation's general coefficient path)."""
    print("case: march material in absorber", flush=True)
    sc = absorber_material_scene(48, 8)
    _same(_run(sc, E.KERNEL_MARCH, 120), _run(sc, E.KERNEL_AUTO, 120))


def test_march_far_field_partial_chunk_and_group():
    """Far-field side columns written by the march where the Huygens box reaches into the
    lattice's partial last z-chunk (nz = 53: planes 48-52) and partial last row group (ny = 38:
    rows 35-37), with dielectric boxes and a thin-wire dipole inside the box: fields and
    far-field phasors bit-identical to the two-pass kernel."""
    print("case: march side columns, 264 x 38 x 53", flush=True)
    sc = Scene(size=(264, 38, 53), pml=1, courant=0.57)
    sc.box((100, 12, 12), (140, 30, 40), eps_r=2.5, sigma=0.01)
    sc.box((20, 8, 8), (60, 20, 20), eps_r=4.0)
    sc.dipole_thin((132, 19, 26), axis=2, length=9.0, radius=0.1, impedance=50.0, amplitude=1.0, wavelength=16.0,
                   waveform="pulse", pulse_width=1.0)
    sc.emit_thin_wires()
    ff = sc.set_far_field(margin=0, wavelength=16.0)
    assert ff.box_max[1] >= 35 and ff.box_max[2] >= 48
    _same(_run(sc, E.KERNEL_MARCH, 100), _run(sc, E.KERNEL_AUTO, 100))


def test_march_far_field_box_through_x_seam():
    """A far-field box whose x faces are maintained as side columns by the march (and by the
    recomputation for material texels on them), on a lattice of two segments."""
    print("case: march side columns, nx 264", flush=True)
    sc = Scene(size=(264, 40, 40), pml=6, courant=0.57)
    sc.box((100, 10, 10), (140, 30, 30), eps_r=2.5, sigma=0.01)
    sc.hertzian((132.0, 20.0, 20.0), (0, 0, 1), 1.0, wavelength=16.0, waveform="pulse", pulse_width=1.0)
    sc.set_far_field(wavelength=16.0)
    _same(_run(sc, E.KERNEL_MARCH, 100), _run(sc, E.KERNEL_AUTO, 100))
