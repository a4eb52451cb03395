"""Headless frames of the volume view (the viewer's raymarch over the grid's own planes):
steps a scene, renders frames with the reference's orbit camera and writes the last one as a
binary PPM (top row first, as image files are stored).

    python -m em_lab_amd.view --size 256 --steps 300 --frames 5 --mode 0 --out frame.ppm
"""
from __future__ import annotations

import argparse
import time

import numpy as np

from em_lab_amd import emfdtd as E
from em_lab_amd.scene import spheres_dipole


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--size", type=int, default=256, help="lattice cells per axis (spheres + dipole scene)")
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--frames", type=int, default=5)
    ap.add_argument("--mode", type=int, default=E.RENDER_FIELDS, help="0 E+B, 1 E, 2 B, 3 Poynting, 4 signed E")
    ap.add_argument("--width", type=int, default=1280)
    ap.add_argument("--height", type=int, default=720)
    ap.add_argument("--out", default="")
    a = ap.parse_args(argv)
    sc = spheres_dipole(n=a.size, pml=10, n_spheres=32, wavelength=20.0)
    g = E.SimulationGrid(0)
    try:
        g.configure(sc.grid_config())
        s, e = sc.flat_materials()
        g.upload_materials(s, e, sc.subcell)
        g.set_sources(sc.sources)
        g.step(a.steps)
        g.synchronize()
        Es, Bs = g.fields(E.LAYOUT_SOA)
        gains = dict(gain_e=2.0 / max(float(np.abs(Es).max()), 1e-30), gain_b=2.0 / max(float(np.abs(Bs).max()), 1e-30))
        cam = E.Camera(distance=3.2, yaw_deg=35.0, pitch_deg=25.0, aspect=a.width / a.height)
        st = E.RenderSettings(mode=a.mode, **gains)
        r = E.VolumeRenderer()
        img = r.render(g, cam, st, a.width, a.height)  # first frame: module load
        t = time.time()
        for _ in range(a.frames):
            img = r.render(g, cam, st, a.width, a.height)
        ms = (time.time() - t) / max(a.frames, 1) * 1e3
        print(f"view: {a.size}^3 after {a.steps} steps, mode {a.mode}: {ms:.2f} ms per {a.width}x{a.height} frame "
              f"(image copied to the host)", flush=True)
        if a.out:
            with open(a.out, "wb") as f:
                f.write(f"P6 {a.width} {a.height} 255\n".encode())
                f.write(np.ascontiguousarray(img[::-1, :, :3]).tobytes())
    finally:
        g.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
