"""Throughput of the FDTD hot path on one MI355X.

A step is one full leapfrog step of the 256^3 scene (32 seeded dielectric spheres, 10-cell
absorber, z thin-wire half-wave dipole behind a 50 ohm resistive feed): fused B+E update,
source injection, port V/I accumulation with series recording, and the near-to-far-field
phasor accumulation on the Huygens box. Prints one JSON line.

value    = whole-job Mcell-updates/s (cells x timed steps / wall time between synchronisations)
roofline = the Yee step (default: the two vectorised passes k_update_b4 + k_update_e4, timed
           together): algorithmic 48 B per cell-update (E and B read once, written once, FP32) /
           its average time from HIP events on the library's stream, bracketing the Yee launches
           of one timed step in TIMING_EVERY (an event pair costs the stream a few microseconds;
           bracketing every step would slow the run it measures)
cpu_baseline = the CPU oracle (scalar C, OpenMP over z-planes) on a bounded number of steps
surrogate = the learned screening path: 65,536 seeded thin-wire antenna candidates through the
           trained MLP (BF16 MFMA), the best 256 by predicted -10 dB bandwidth verified by the
           batched FDTD stepper (4000 steps each); designs screened per second against
           one-at-a-time FDTD of the same designs, and how many verified within 0.5 dB of S11
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

TIMING_EVERY = 16

METRIC = "Mcell-updates/s & %HBM peak; designs/s vs plain FDTD at |\u0394S11|\u22640.5 dB"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = 48  # FP32 E and B (6 components), each read once and written once


def cpu_baseline(scene, budget_s: float = 12.0, max_steps: int = 40):
    """Time the oracle's step loop (update_B, update_E, sources) on the host cores."""
    import numpy as np
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    cfg = scene.grid_config()
    sim = O.OracleSim(cfg.size, cfg.courant, cfg.pml_thickness, cfg.pml_sigma_max, cfg.pml_grading)
    s, e = scene.flat_materials()
    sim.upload_materials(s, e, scene.subcell.edges, scene.subcell.wire_terms)
    sim.set_sources(scene.sources)
    sim.step(1)  # first touch of the arrays
    n = int(np.prod(cfg.size))
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps:
        sim.step(1)
        steps += 1
        if time.perf_counter() - t0 > budget_s and steps >= 3:
            break
    dt = time.perf_counter() - t0
    return {"value": round(n * steps / dt / 1e6, 3), "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{cfg.size[0]}x{cfg.size[1]}x{cfg.size[2]} same scene, {steps} steps of the oracle's "
                      f"update_B/update_E/apply_sources (OpenMP {threads} threads), {dt:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--kernel", type=int, default=0, help="EMFD_KERNEL_* variant (| tile << 8 | zchunk << 16)")
    ap.add_argument("--no-monitors", action="store_true")
    ap.add_argument("--scene", choices=["spheres", "vacuum", "array"], default="spheres",
                    help="spheres: the headline workload; vacuum: same grid, no materials (roofline check); "
                         "array: seeded spheres around a 4x4 thin-wire dipole array (the precision-form scene)")
    ap.add_argument("--storage", choices=["f32", "f16", "bf16"], default="f32",
                    help="field storage: FP32 planes, or 16-bit planes with FP32 updates")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the learned screening measurement (designs/s against one-at-a-time FDTD)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="only time a plain float4 device copy of MB megabytes (roofline calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the single-GPU figure is the per-GPU share; see the sharding note in DESIGN.md)", file=sys.stderr)
        return 2

    from em_lab_amd import emfdtd as E
    from em_lab_amd.scene import hertzian_vacuum, spheres_array, spheres_dipole

    if args.copy_probe:
        g = E.SimulationGrid(0)
        nbytes = args.copy_probe * 1000 * 1000
        ms = g.copy_probe(nbytes, max(args.steps, 1))
        print(json.dumps({"copy_probe_bytes": nbytes, "avg_ms": round(ms, 5),
                          "read_plus_write_GBs": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}), flush=True)
        g.close()
        return 0

    if args.scene == "spheres":
        scene = spheres_dipole(n=args.size, pml=10)
    elif args.scene == "array":
        scene = spheres_array(n=args.size, pml=10)
    else:
        scene = hertzian_vacuum(args.size, 10, steps_wavelength=20.0)
    g = E.SimulationGrid(0)
    g.configure(scene.grid_config())
    s, e = scene.flat_materials()
    g.upload_materials(s, e, scene.subcell)
    g.set_sources(scene.sources)
    g.set_kernel(args.kernel)
    storage = {"f32": E.STORAGE_F32, "f16": E.STORAGE_F16, "bf16": E.STORAGE_BF16}[args.storage]
    if storage != E.STORAGE_F32:
        # E near the feeds peaks at O(1); B at O(1/8): stored values stay far below 65504
        g.set_storage(storage, [0, 0, 0, -3, -3, -3])
    bytes_per_cell = BYTES_PER_CELL if storage == E.STORAGE_F32 else BYTES_PER_CELL // 2
    if not args.no_monitors:
        if scene.ports:
            pm = E.PortMonitor(g)
            pm.configure(scene.ports, scene.far_field.averaging_periods, scene.sweep, scene.cell_size)
        ff = E.FarFieldMonitor(g)
        ff.configure(scene.far_field)
        # the workload's far field is evaluated at three wavelengths (lambda / 1.333, lambda, 1.5 lambda)
        wl = scene.far_field.wavelength
        ff.set_wavelengths([wl, wl * 0.75, wl * 1.5])
    nx, ny, nz = g.size
    cells = nx * ny * nz

    g.step(args.warmup)
    g.synchronize()
    g.kernel_timing(TIMING_EVERY)
    t0 = time.perf_counter()
    g.step(args.steps)
    g.synchronize()
    t1 = time.perf_counter()
    yee_ms, launches = g.kernel_timing(0)
    wall = t1 - t0
    Eg, _ = g.fields(E.LAYOUT_SOA)
    import numpy as np
    if not np.all(np.isfinite(Eg)):
        print("bench.py: non-finite fields", file=sys.stderr)
        return 2

    value = cells * args.steps / wall / 1e6
    pattern = None
    if not args.no_monitors:
        # far-field pattern of the workload: a 181 x 361 (theta, phi) grid at the three wavelengths
        dirs = E.FarFieldMonitor.theta_phi_grid(181, 361)
        ff.radiation(dirs[:256], 0)  # warm-up
        t_p = time.perf_counter()
        peaks = []
        for f in range(3):
            f2, P = ff.radiation(dirs, f)
            peaks.append(float(f2.max()))
        dt_p = time.perf_counter() - t_p
        pairs = 3 * dirs.shape[0] * ff.n_points
        pattern = {"directions": int(dirs.shape[0]), "points": int(ff.n_points), "wavelengths": 3,
                   "seconds": round(dt_p, 4), "pairs_per_s": round(pairs / dt_p, 1),
                   "flop_per_pair": "sincos + 12 FMA"}
    avg_ms = yee_ms / max(launches, 1)
    achieved = bytes_per_cell * cells / (avg_ms * 1e-3) / 1e9
    # HBM bytes per Yee step from the calibrated rocprofv3 counter runs of this workload
    # (profiles/yee_256_traffic.md); only quoted for the default kernel at 256^3
    traffic = None
    tf = ROOT / "profiles" / "yee_256_traffic.json"
    if (tf.exists() and args.size == 256 and args.kernel in (0, 5) and args.scene == "spheres"
            and storage == E.STORAGE_F32):
        traffic = json.loads(tf.read_text()).get("bytes_per_launch")
    kernel_names = {0: "k_update_b4+k_update_e4", 1: "k_yee_fused", 2: "k_update_b+k_update_e", 3: "k_yee_rows",
                    4: "k_yee_point", 5: "k_update_b4+k_update_e4", 6: "k_update_b4+k_update_e4 (z-slabs)", 7: "k_yee_zmarch"}
    kname = (kernel_names.get(args.kernel & 0xff, "?") if storage == E.STORAGE_F32
             else "k_update_bh+k_update_eh (%d cells per thread)" % ((args.kernel >> 8) & 0xf or 8))
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (f"{nx}^3 Yee leapfrog, 10-cell absorber, 32 seeded dielectric spheres, z thin-wire "
                                f"half-wave dipole (lambda 20) on a 50 ohm resistive feed, port (32-point sweep) "
                                f"+ far-field (3 wavelengths) accumulation every step" if args.scene == "spheres" else
                                f"{nx}^3 Yee leapfrog, 10-cell absorber, 32 seeded dielectric spheres around a 4x4 "
                                f"thin-wire dipole array on 50 ohm feeds, port + far-field accumulation every step"
                                if args.scene == "array" else
                                f"{nx}^3 vacuum Yee leapfrog, 10-cell absorber, z Hertzian pulse, far-field "
                                f"accumulation every step") + (" (monitors off)" if args.no_monitors else "")
                               + ("" if args.storage == "f32" else f", {args.storage} field storage"),
                   "grid": [nx, ny, nz], "courant": round(g.dt, 6), "material_classes": g.material_classes(),
                   "kernel": args.kernel, "field_storage": args.storage},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kname, "avg_launch_ms": round(avg_ms, 5),
                     "launches": launches, "timed_every": TIMING_EVERY,
                     "bytes_per_cell": bytes_per_cell},
        "cpu_baseline": None,
    }
    if pattern is not None:
        out["ntff_pattern"] = pattern
    g.close()
    if (not args.no_surrogate and args.size == 256 and args.scene == "spheres" and args.storage == "f32"
            and args.kernel == 0):
        from em_lab_amd import surrogate as S
        out["surrogate"] = S.screening_run(65536, 256, plain_designs=4)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scene)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
