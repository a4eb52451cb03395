"""Coax line monitor on the GPU (emfd_lines_*, k_line_accumulate) against the oracle's
or_line_accumulate (shaders/line_accumulate.comp:30-48) stepping the same scene in lockstep:
phasors and instantaneous values bit for bit after every step call, FP32 and binary16 storage.
The lines sit beside the Hertzian source so that their fields are not zero; one line points
down -z and one runs along x through a balun sleeve, and some terms reach outside the grid
(clamped)."""
import numpy as np
import pytest

from em_lab_amd import coax as X
from em_lab_amd import emfdtd as E
from em_lab_amd.scene import hertzian_vacuum

from gpu_helpers import gpu_grid, oracle_sim

pytestmark = pytest.mark.gpu


def _lines(n):
    return [X.CoaxSetup(name="z", axis=2, sign=-1, start=n - 6, center=(n / 2 + 4.5, n / 2 + 0.5, 0.0), length=24,
                        inner=1, outer=5, wavelength=16.0),
            X.CoaxSetup(name="x", axis=0, sign=1, start=-3, center=(0.0, n / 2 - 3.5, n / 2 + 0.5), length=n,
                        inner=1, outer=5, sleeve_start=6, sleeve_end=14, sleeve_outer=11, wavelength=16.0)]


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_lines_bit_exact(storage):
    n = 40
    print(f"case: coax line monitor {storage} {n}^3, 2 lines, 3 calls of 7 steps", flush=True)
    scene = hertzian_vacuum(n, 6)
    g, _, _ = gpu_grid(scene, ports=False, ntff=False)
    sim = oracle_sim(scene, ports=False, ntff=False)
    if storage == "f16":
        k = [0, 0, 0, -3, -3, -3]
        g.set_storage(E.STORAGE_F16, k)
        sim.set_storage(1, k)
    lines = _lines(n)
    mon = X.CoaxMonitor(g)
    mon.configure(lines, 3.0)
    q, t, w, _ = X.build_terms(lines)
    assert t[:, :3].min() < 0 or t[:, :3].max() >= n  # some terms clamp
    sim.configure_lines(q, t, w, lines[0].wavelength, 3.0)
    for call in range(3):
        g.step(7)
        sim.step(7)
        ph, now = mon._dev.read()
        assert np.array_equal(now, sim.lines["now"]), f"call {call}: instantaneous values differ"
        assert np.array_equal(ph, sim.lines["phasors"]), f"call {call}: phasors differ"
    assert np.abs(sim.lines["now"]).max() > 0.0
    r = mon.readings()
    assert len(r) == 2 and len(r[0].V) == 23 and len(r[1].V) == n - 1
    g.close()


def test_lines_configure_errors_and_disable():
    scene = hertzian_vacuum(24, 4)
    g, _, _ = gpu_grid(scene, ports=False, ntff=False)
    dev = E.LineMonitor(g)
    with pytest.raises(E.EmfdError):
        dev.configure([[0, 3, 0, 0]], [[1, 1, 1, 0]] * 2, [1.0, 1.0], 16.0)  # past the term list
    with pytest.raises(E.EmfdError):
        dev.configure([[0, 1, 2, 0]], [[1, 1, 1, 0]], [1.0], 16.0)  # unknown kind
    with pytest.raises(E.EmfdError):
        dev.configure([[0, 1, 0, 0]], [[1, 1, 1, 6]], [1.0], 16.0)  # component outside 0..5
    dev.configure([[0, 1, 0, 0]], [[12, 12, 12, 2]], [1.0], 16.0)
    g.step(20)
    ph, now = dev.read()
    assert ph.shape == (1, 4) and now[0] != 0.0
    dev.configure([], [], [], 16.0)  # disabled: steps run without the monitor
    g.step(2)
    assert dev.read()[0].shape == (0, 4)
    g.close()
