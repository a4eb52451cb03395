"""16-bit field storage on the GPU against the oracle's storage emulation: every pass's output
rounded to binary16 / bfloat16 of value / 2^k; all arithmetic FP32. Bit-exact, like the FP32
path, for the two-pass kernels (8 or 4 cells per thread) and for the one-pass march, which
rounds B^{n+1/2} through the storage before the E update reads it."""
import numpy as np
import pytest

from em_lab_amd import emfdtd as E
from em_lab_amd.scene import hertzian_vacuum, spheres_dipole

from gpu_helpers import absorber_material_scene, gpu_grid, oracle_sim, rel_l2

pytestmark = pytest.mark.gpu


KERNELS16 = [E.KERNEL_AUTO, E.KERNEL_AUTO | (4 << 8), E.KERNEL_MARCH, E.KERNEL_MARCH | (3 << 8)]


@pytest.mark.parametrize("kernel", KERNELS16)
@pytest.mark.parametrize("storage", [E.STORAGE_F16, E.STORAGE_BF16])
def test_hertzian_storage_bit_exact(storage, kernel):
    """kernel: the two-pass kernels at 8 or 4 cells per thread (bits 8.. of the automatic
    variant), or the march (bits 8..: z-chunk, 0 = 8)."""
    print("case: hertzian 64^3 storage", storage, "kernel", kernel, flush=True)
    sc = hertzian_vacuum(64, 8)
    k = [2, 2, 2, -3, -3, -3]
    g, _, ff = gpu_grid(sc, kernel)
    g.set_storage(storage, k)
    sim = oracle_sim(sc, gpu_ntff_setup=ff.setup)
    sim.set_storage(storage, k)
    g.step(200)
    sim.step(200)
    Eg, Bg = g.fields(E.LAYOUT_TEXEL)
    assert np.array_equal(Eg[:, :3], sim.E[:, :3]) and np.array_equal(Bg[:, :3], sim.B[:, :3])
    assert np.any(Eg[:, :3])
    assert np.array_equal(ff.phasors(), sim.ntff["phasors"])
    g.close()


@pytest.mark.parametrize("kernel", [E.KERNEL_AUTO, E.KERNEL_MARCH])
@pytest.mark.parametrize("storage", [E.STORAGE_F32, E.STORAGE_F16, E.STORAGE_BF16])
def test_far_field_box_through_source(storage, kernel):
    """A far-field box whose upper x face (x = 32) runs through the Hertzian source at the centre
    of a 64^3 grid: the x-face side columns the Yee and source passes keep must hold exactly what
    the planes hold (the stored, rounded values for 16-bit storage), so the phasors stay equal to
    the oracle's bit for bit."""
    print("case: far-field box through the source, storage", storage, "kernel", kernel, flush=True)
    sc = hertzian_vacuum(64, 8)
    ff0 = sc.far_field
    sc.far_field = E.FarFieldSetup((14, 13, 12), (32, 50, 49), ff0.wavelength, ff0.averaging_periods, ff0.map_axis)
    k = [2, 2, 2, -3, -3, -3]
    g, _, ff = gpu_grid(sc, kernel)
    sim = oracle_sim(sc, gpu_ntff_setup=ff.setup)
    if storage != E.STORAGE_F32:
        g.set_storage(storage, k)
        sim.set_storage(storage, k)
    g.step(120)
    sim.step(120)
    Eg, Bg = g.fields(E.LAYOUT_TEXEL)
    assert np.array_equal(Eg[:, :3], sim.E[:, :3]) and np.array_equal(Bg[:, :3], sim.B[:, :3])
    ph = ff.phasors()
    assert np.abs(ph).max() > 0
    assert np.array_equal(ph, sim.ntff["phasors"])
    g.close()


@pytest.mark.parametrize("kernel", [E.KERNEL_AUTO, E.KERNEL_MARCH])
@pytest.mark.parametrize("storage", [E.STORAGE_F16, E.STORAGE_BF16])
def test_material_in_absorber_storage(storage, kernel):
    print("case: material in the absorber 48^3, storage", storage, "kernel", kernel, flush=True)
    sc = absorber_material_scene()
    k = [2, 2, 2, -3, -3, -3]
    g, _, _ = gpu_grid(sc, kernel, ntff=False, ports=False)
    g.set_storage(storage, k)
    sim = oracle_sim(sc, ntff=False, ports=False)
    sim.set_storage(storage, k)
    g.step(80)
    sim.step(80)
    Eg, Bg = g.fields(E.LAYOUT_TEXEL)
    assert np.array_equal(Eg[:, :3], sim.E[:, :3]) and np.array_equal(Bg[:, :3], sim.B[:, :3])
    assert np.any(Eg[:, :3])
    g.close()


@pytest.mark.parametrize("kernel", [E.KERNEL_AUTO, E.KERNEL_MARCH])
@pytest.mark.parametrize("storage", [E.STORAGE_F16, E.STORAGE_BF16])
def test_thick_absorber_partial_rows_storage(storage, kernel):
    """16-bit kernels with a 36-cell absorber (two rows of 74 depth classes staged per wavefront:
    more than one pass over the lanes) on a 96-wide lattice (a row fills 12 of 32 lanes) and
    dielectric boxes reaching into the absorber: bit for bit against the oracle."""
    print("case: 96^3, 36-cell absorber, storage", storage, "kernel", kernel, flush=True)
    sc = absorber_material_scene(n=96, pml=36)
    k = [2, 2, 2, -3, -3, -3]
    g, _, _ = gpu_grid(sc, kernel, ntff=False, ports=False)
    g.set_storage(storage, k)
    sim = oracle_sim(sc, ntff=False, ports=False)
    sim.set_storage(storage, k)
    g.step(30)
    sim.step(30)
    Eg, Bg = g.fields(E.LAYOUT_TEXEL)
    assert np.array_equal(Eg[:, :3], sim.E[:, :3]) and np.array_equal(Bg[:, :3], sim.B[:, :3])
    assert np.any(Eg[:, :3])
    g.close()


@pytest.mark.parametrize("kernel", KERNELS16[:3])
def test_spheres_dipole_f16_ports_bit_exact(kernel):
    print("case: spheres+dipole 48^3 f16 storage with ports, kernel", kernel, flush=True)
    sc = spheres_dipole(n=48, pml=8, n_spheres=4, radius_range=(3.0, 6.0), sweep_points=16)
    k = [0, 0, 0, -4, -4, -4]
    g, pm, _ = gpu_grid(sc, kernel, ntff=False)
    g.set_storage(E.STORAGE_F16, k)
    sim = oracle_sim(sc, ntff=False)
    sim.set_storage(1, k)
    g.step(150)
    sim.step(150)
    Eg, Bg = g.fields(E.LAYOUT_TEXEL)
    assert np.array_equal(Eg[:, :3], sim.E[:, :3]) and np.array_equal(Bg[:, :3], sim.B[:, :3])
    assert np.array_equal(pm.phasors().ravel(), sim.ports["phasors"].ravel()[:8 * len(sc.ports)])
    # the 16-bit run stays close to the FP32 run of the same scene
    g32, _, _ = gpu_grid(sc, ntff=False)
    g32.step(150)
    E32, _ = g32.fields(E.LAYOUT_TEXEL)
    assert rel_l2(Eg[:, :3], E32[:, :3]) < 2e-2
    g.close()
    g32.close()


def test_storage_roundtrip_and_errors():
    sc = hertzian_vacuum(32, 4)
    g, _, _ = gpu_grid(sc, ntff=False)
    g.set_storage(E.STORAGE_F16, [1, 0, -1, 0, 0, 0])
    n = 32 ** 3
    rng = np.random.default_rng(0)
    e0 = rng.standard_normal((3, n)).astype(np.float32)
    g.set_fields(e0, None, E.LAYOUT_SOA)
    e1, _ = g.fields(E.LAYOUT_SOA)
    scale = np.array([2.0, 1.0, 0.5], np.float32)[:, None]
    assert np.array_equal(e1, (e0 / scale).astype(np.float16).astype(np.float32) * scale)
    with pytest.raises(E.EmfdError):
        g.set_kernel(E.KERNEL_TWO_PASS4)
        g.step(1)
    g.set_kernel(E.KERNEL_AUTO)
    g.step(2)
    g.close()


def test_f64_stepper_against_fp32_oracle():
    """The FP64 stepper computes the same model in double: it agrees with the FP32 oracle to FP32
    rounding on a scene with every ingredient (absorber, spheres, thin wire, resistive feed)."""
    print("case: f64 stepper spheres+dipole 48^3", flush=True)
    sc = spheres_dipole(n=48, pml=8, n_spheres=4, radius_range=(3.0, 6.0), sweep_points=16)
    g, pm, _ = gpu_grid(sc, ntff=False)
    g.set_storage(E.STORAGE_F64)
    sim = oracle_sim(sc, ntff=False)
    g.step(200)
    sim.step(200)
    Ed, Bd = g.fields_f64()
    Eo = sim.E[:, :3].T.astype(np.float64)
    Bo = sim.B[:, :3].T.astype(np.float64)
    assert rel_l2(Ed, Eo) <= 1e-5 and rel_l2(Bd, Bo) <= 1e-5, (rel_l2(Ed, Eo), rel_l2(Bd, Bo))
    assert not np.array_equal(Ed, Eo)  # really computed in double
    assert rel_l2(pm.phasors(), sim.ports["phasors"][:len(sc.ports)]) <= 1e-4
    g.close()


def test_f64_cavity_energy_invariant():
    """pml = 0: the leapfrog invariant |E^n|^2 + <B^{n+1/2}, B^{n-1/2}> holds to FP64 rounding."""
    from em_lab_amd.scene import Scene
    print("case: f64 cavity energy", flush=True)
    sc = Scene(size=(40, 36, 30), pml=0, courant=0.5, pml_sigma_max=0.0)
    g, _, _ = gpu_grid(sc, ports=False, ntff=False)
    g.set_storage(E.STORAGE_F64)
    rng = np.random.default_rng(8)
    n = 40 * 36 * 30
    g.set_fields(rng.standard_normal((3, n)).astype(np.float32), rng.standard_normal((3, n)).astype(np.float32),
                 E.LAYOUT_SOA)
    inv = []
    for _ in range(12):
        e0, b0 = g.fields_f64()
        g.step(1)
        _, b1 = g.fields_f64()
        inv.append(np.sum(e0 ** 2) + np.sum(b1 * b0))
    inv = np.array(inv)
    assert np.max(np.abs(inv - inv[0])) / abs(inv[0]) < 1e-12
    g.close()
