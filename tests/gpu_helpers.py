"""Shared drivers for the GPU parity tests: run one scene on the HIP library and on the CPU
oracle with the same inputs, in the reference's call order."""
from __future__ import annotations

import numpy as np

from em_lab_amd import emfdtd as E
from em_lab_amd.scene import Scene
from oracle import oracle as O


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def gpu_grid(scene: Scene, kernel: int = E.KERNEL_AUTO, ports=True, ntff=True, record=False):
    g = E.SimulationGrid(0)
    g.configure(scene.grid_config())
    s, e = scene.flat_materials()
    g.upload_materials(s, e, scene.subcell)
    g.set_sources(scene.sources)
    g.set_kernel(kernel)
    pm = ff = None
    if ports and scene.ports:
        pm = E.PortMonitor(g)
        avg = scene.far_field.averaging_periods if scene.far_field else 4.0
        pm.configure(scene.ports, avg, scene.sweep, scene.cell_size)
    if ntff and scene.far_field:
        ff = E.FarFieldMonitor(g)
        ff.configure(scene.far_field)
    return g, pm, ff


def oracle_sim(scene: Scene, ports=True, ntff=True, gpu_ntff_setup=None, capacity=0):
    cfg = scene.grid_config()
    sim = O.OracleSim(cfg.size, cfg.courant, cfg.pml_thickness, cfg.pml_sigma_max, cfg.pml_grading)
    s, e = scene.flat_materials()
    sim.upload_materials(s, e, scene.subcell.edges, scene.subcell.wire_terms)
    sim.set_sources(scene.sources)
    if ports and scene.ports:
        avg = scene.far_field.averaging_periods if scene.far_field else 4.0
        sim.configure_ports(scene.ports, avg, record=scene.sweep is not None, capacity=capacity)
    if ntff and scene.far_field:
        st = gpu_ntff_setup or scene.far_field
        sim.configure_ntff(st.box_min, st.box_max, st.wavelength, st.averaging_periods)
    return sim


def absorber_material_scene(n: int = 48, pml: int = 8):
    """Hertzian n^3 vacuum scene with two dielectric boxes that reach into the absorber (one lossy
    into the lower x layer, one through both y layers): the E pass's material-in-absorber path."""
    from em_lab_amd.scene import hertzian_vacuum
    sc = hertzian_vacuum(n, pml)
    sc.box((2, 10, 10), (20, 30, 30), eps_r=3.0, sigma=0.05)
    sc.box((26, 2, 20), (30, n - 2, 24), eps_r=2.0)
    return sc
