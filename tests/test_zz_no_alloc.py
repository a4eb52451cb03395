"""emfd_flux_measure_noalloc through the C ABI: the same sums as emfd_flux_measure, bit for bit,
on two calls in a row (the second with fewer and smaller surfaces, so the device array holds the
first call's larger contents when it starts), and its limits rejected on the host."""
import ctypes as C

import numpy as np
import pytest

from em_lab_amd import emfdtd as E
from em_lab_amd.scene import spheres_dipole

from gpu_helpers import gpu_grid

pytestmark = pytest.mark.gpu


def _measure(g, fn, surfaces):
    arr = (E.FluxSurfaceC * max(len(surfaces), 1))(*[s.to_c() for s in surfaces])
    out = np.full(max(len(surfaces), 1), np.nan, np.float64)
    rc = getattr(g._lib, fn)(g.handle, C.cast(arr, C.c_void_p), len(surfaces), out.ctypes.data_as(C.c_void_p))
    return rc, out[: len(surfaces)]


def _both(g, surfaces):
    ra, a = _measure(g, "emfd_flux_measure", surfaces)
    rb, b = _measure(g, "emfd_flux_measure_noalloc", surfaces)
    assert ra == 0 and rb == 0
    return a, b


def test_flux_noalloc_matches_flux_measure():
    print("case: no-allocation flux vs emfd_flux_measure, 64^3 spheres + dipole", flush=True)
    sc = spheres_dipole(n=64, pml=8, n_spheres=4, wavelength=16.0, radius_range=(4.0, 8.0), sweep_points=16)
    g, _, _ = gpu_grid(sc, ports=False, ntff=False)
    F = E.FluxSurface
    g.step(90)
    first = [F.box((12, 12, 12), (52, 52, 52)), F.box((30, 20, 25), (20, 44, 40)), F.box((0, 0, 0), (63, 63, 63)),
             F.box((-5, 10, 10), (70, 50, 30))]
    first += [F.plane(a, i, (4, 6), (60, 58)) for a in range(3) for i in (9, 32, 55)]
    first += [F.plane(0, 80, (0, 0), (64, 64)), F.plane(1, 20, (70, 0), (90, 10)), F.plane(2, 31, (-8, -8), (100, 100))]
    a, b = _both(g, first)
    print("first call:", a[:4], flush=True)
    assert np.array_equal(a, b)
    assert np.abs(a).max() > 0
    assert a[len(first) - 3] == 0.0 and a[len(first) - 2] == 0.0  # planes outside the lattice
    g.step(25)
    second = [F.box((28, 28, 28), (36, 36, 36)), F.plane(2, 40, (20, 20), (44, 44))]
    a, b = _both(g, second)
    print("second call:", a, flush=True)
    assert np.array_equal(a, b)
    assert np.abs(a).max() > 0
    g.close()


def test_flux_noalloc_limits():
    print("case: no-allocation flux limits", flush=True)
    sc = spheres_dipole(n=32, pml=4, n_spheres=1, wavelength=12.0, radius_range=(2.0, 3.0), sweep_points=8)
    g, _, _ = gpu_grid(sc, ports=False, ntff=False)
    F = E.FluxSurface
    rc, _ = _measure(g, "emfd_flux_measure_noalloc", [F.box((4, 4, 4), (20, 20, 20))] * 65)
    assert rc == E.EMFD_ERR_INVALID
    rc, out = _measure(g, "emfd_flux_measure_noalloc", [])
    assert rc == 0 and out.size == 0
    g.close()
    # 64 closed boxes over a 16 x 16 x 1100 lattice: 64 * (4 * 1100 + 2 * 16) rows > 262144
    from em_lab_amd.scene import Scene
    long = Scene(size=(16, 16, 1100), pml=2)
    g, _, _ = gpu_grid(long, ports=False, ntff=False)
    rc, _ = _measure(g, "emfd_flux_measure_noalloc", [F.box((0, 0, 0), (15, 15, 1099))] * 64)
    assert rc == E.EMFD_ERR_INVALID
    rc, a = _measure(g, "emfd_flux_measure_noalloc", [F.box((0, 0, 0), (15, 15, 1099))] * 59)
    assert rc == 0 and np.array_equal(a, np.zeros(59))
    g.close()
