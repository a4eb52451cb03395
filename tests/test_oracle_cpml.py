"""The oracle's convolutional PML (or_cpml_*, the library's opt-in absorber) against the closed
form of a stretched-coordinate layer: the reflection of a plane pulse at normal incidence, with
the graded profile sigma = sigma_max d^3 over an 8-cell layer.

A lattice of this thickness samples the profile coarsely: the measured attenuation exceeds the
continuous integral by 10-20% over 8-16 cells per wavelength (it falls toward it as the layer
thickens; the GPU tests use 16 cells). The bound below allows that and no more: a dropped
factor in the auxiliary recursion, a wrong sign of its correction or a profile on the wrong
stagger moves ln R by far more. Parity unpinned beyond that bound for this function."""
import numpy as np

from em_lab_amd.scene import MAX_COURANT
from oracle import oracle as O

from cpml_reflection import closed_ln_r, measured_r, record_port, sheet_sources

L, SIGMA_MAX, GRADING = 8, 0.5, 3.0
LAM, PW = 8.0, 0.35


def _run(nx, n_t, xs, xo, steps, cpml):
    sim = O.OracleSim((nx, n_t, n_t), MAX_COURANT, L, 1.6, 3.0)
    sim.set_sources(sheet_sources(n_t, n_t, xs, LAM, PW, 0.0, 1))
    if cpml:
        sim.set_cpml(SIGMA_MAX, GRADING, 0.0)
    sim.configure_ports([record_port((xo, n_t // 2, n_t // 2), 1, LAM)], 4.0, record=True, capacity=steps)
    sim.step(steps)
    return sim.ports["series"][:steps, 0, 0].copy()


def test_oracle_cpml_normal_reflection():
    xs = L + 16
    xo = xs + 5
    nx = xo + 8 + L + 1
    n_t, steps = 120, 100
    vt = _run(nx, n_t, xs, xo, steps, True)
    vr = _run(nx + 40, n_t, xs, xo, steps, True)
    lams = [10.0, 12.0, 16.0]
    r = measured_r(vt, vr, MAX_COURANT, 30.0, lams)
    ratio = np.array([np.log(v) / closed_ln_r(SIGMA_MAX, GRADING, L, lam) for v, lam in zip(r, lams)])
    print("R", r, "ln R / closed form", ratio)
    assert np.all(np.abs(vt[: int(28 / MAX_COURANT)] - vr[: int(28 / MAX_COURANT)]) < 1e-4)  # same until it returns
    assert np.all(ratio > 0.95) and np.all(ratio < 1.25)
