"""Context lifetime: emfd_destroy returns every device buffer the context allocated, including the
march's material tables and sparse lists, the far-field sample slots and the render staging
image, so a process that builds and drops grids (a sweep driver, the test suite) does not grow."""
import ctypes as C

import numpy as np
import pytest

from em_lab_amd import emfdtd as E
from em_lab_amd.scene import spheres_dipole

pytestmark = pytest.mark.gpu


def _free_bytes() -> int:
    h = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert h.hipDeviceSynchronize() == 0
    assert h.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _cycle(sc):
    g = E.SimulationGrid(0)
    g.configure(sc.grid_config())
    s, e = sc.flat_materials()
    g.upload_materials(s, e, sc.subcell)
    g.set_sources(sc.sources)
    ff = E.FarFieldMonitor(g)
    ff.configure(sc.far_field)
    pm = E.PortMonitor(g)
    pm.configure(sc.ports, sc.far_field.averaging_periods, sc.sweep, sc.cell_size)
    g.step(20)
    E.VolumeRenderer().render(g, E.Camera(), E.RenderSettings(), 64, 36)
    g.synchronize()
    assert g.material_classes() > 0
    g.close()


def test_destroy_releases_device_memory():
    print("case: 3 create/step/destroy cycles of a 256^3 material scene (march tables)", flush=True)
    sc = spheres_dipole(n=256, pml=10, n_spheres=8, wavelength=20.0)
    _cycle(sc)  # first use: runtime and module state
    before = _free_bytes()
    for _ in range(3):
        _cycle(sc)
    after = _free_bytes()
    # one leaked set of march class bytes alone would be 16.8 MB per cycle
    assert before - after < 8 << 20, (before - after) / 2**20
