"""Throughput of the FDTD hot path on one MI355X.

Headline workload: the precision-form scene at 768^3 in FP32 (32 seeded dielectric spheres around
a 4 x 4 array of z thin-wire half-wave dipoles on 50 ohm resistive feeds, 10-cell absorber), with
the port monitor (16 ports, sweep recording) and the far-field accumulation at three wavelengths
running every step. A step is one full leapfrog step: B and E update, source injection, port V/I
accumulation, far-field phasor accumulation. Prints one JSON line.

value    = whole-job Mcell-updates/s (cells x timed steps / wall time between synchronisations)
roofline = the Yee update: algorithmic 48 B per cell-update (E and B read once, written once, FP32)
           / its median launch time from HIP events on the library's stream around the Yee
           launches of the timed steps (at least 8 of them); traffic = HBM bytes per Yee step from
           the committed rocprofv3 FETCH_SIZE / WRITE_SIZE profile of the same workload, or null
           when no profile of that workload exists
cpu_baseline = the CPU oracle (scalar C, OpenMP over z-planes) on the 192^3 copy of the same
           scene: median of 3 runs
extras   = the pure-vacuum 768^3 Yee update, the same scene with the convolutional PML in the
           absorber slabs (its roofline counting the auxiliary fields' bytes too), the same scene
           with binary16 field storage, and the 256^3 sphere / thin-wire dipole scene
surrogate = the learned screening path: 65,536 seeded thin-wire antenna candidates through the
           trained MLP (BF16 MFMA), the best 256 by predicted -10 dB bandwidth verified by the
           batched FDTD stepper (4000 steps each); designs screened per second against
           one-at-a-time FDTD of the same designs, and the surrogate's pick precision
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mcell-updates/s & %HBM peak; designs/s vs plain FDTD at |ΔS11|≤0.5 dB"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = 48  # FP32 E and B (6 components), each read once and written once
# convolutional PML: per texel of an axis's slab (2 pml + 1 planes), its four auxiliary fields
# (B_v, B_w after Faraday, E_v, E_w after Ampere) read once and written once
CPML_AUX_BYTES = 4 * 4 * 2
MIN_TIMED_LAUNCHES = 8


def cpu_info():
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def cpu_baseline(scene_name: str, n: int = 192, steps: int = 30, repeats: int = 3):
    """Median of `repeats` timings of the oracle's step loop (update_B, update_E, sources) on the
    n^3 copy of the scene, on the host cores (OMP_NUM_THREADS, else every core)."""
    from em_lab_amd.scene import spheres_array, spheres_dipole
    from oracle import oracle as O

    model, nproc = cpu_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or nproc
    os.environ["OMP_NUM_THREADS"] = str(threads)
    scene = spheres_array(n=n, pml=10) if scene_name == "array" else spheres_dipole(n=n, pml=10)
    cfg = scene.grid_config()
    sim = O.OracleSim(cfg.size, cfg.courant, cfg.pml_thickness, cfg.pml_sigma_max, cfg.pml_grading)
    s, e = scene.flat_materials()
    sim.upload_materials(s, e, scene.subcell.edges, scene.subcell.wire_terms)
    sim.set_sources(scene.sources)
    sim.step(1)  # first touch of the arrays
    cells = cfg.size[0] * cfg.size[1] * cfg.size[2]
    rates = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        sim.step(steps)
        rates.append(cells * steps / (time.perf_counter() - t0) / 1e6)
    return {"value": round(statistics.median(rates), 3), "unit": "Mcell-updates/s", "cores": threads,
            "kind": "port", "nproc": nproc, "cpu_model": model,
            "runs": [round(r, 3) for r in rates],
            "sample": f"{n}^3 copy of the same scene, median of {repeats} runs of {steps} steps of the oracle's "
                      f"update_B/update_E/apply_sources (OpenMP, {threads} threads)"}


def build_scene(name: str, n: int):
    from em_lab_amd.scene import hertzian_vacuum, spheres_array, spheres_dipole
    if name == "array":
        return spheres_array(n=n, pml=10)
    if name == "spheres":
        return spheres_dipole(n=n, pml=10)
    return hertzian_vacuum(n, 10, steps_wavelength=20.0)


def workload_text(name: str, n: int, storage: str, monitors: bool) -> str:
    base = {"array": f"{n}^3 Yee leapfrog, 10-cell absorber, 32 seeded dielectric spheres around a 4x4 thin-wire "
                     f"dipole array on 50 ohm feeds, port (16 ports, sweep recording) + far-field (3 wavelengths) "
                     f"accumulation every step",
            "spheres": f"{n}^3 Yee leapfrog, 10-cell absorber, 32 seeded dielectric spheres, z thin-wire half-wave "
                       f"dipole (lambda 20) on a 50 ohm resistive feed, port (32-point sweep) + far-field "
                       f"(3 wavelengths) accumulation every step",
            "vacuum": f"{n}^3 vacuum Yee leapfrog, 10-cell absorber, z Hertzian pulse"}[name]
    if not monitors and name != "vacuum":
        base += " (monitors off)"
    return base + ("" if storage == "f32" else f", {storage} field storage")


def traffic_for(key: str):
    """HBM bytes per Yee step of this workload from the committed counter profile, or None."""
    tf = ROOT / "profiles" / "yee_traffic.json"
    if not tf.exists():
        return None
    entry = json.loads(tf.read_text()).get("workloads", {}).get(key)
    return None if entry is None else entry.get("bytes_per_launch")


def run_grid(name: str, n: int, storage: str, kernel: int, steps: int, warmup: int, monitors: bool,
             pattern: bool = False, absorber: str = "graded"):
    """Runs one workload; returns (result dict, whether the fields stayed finite)."""
    import numpy as np
    from em_lab_amd import emfdtd as E

    scene = build_scene(name, n)
    g = E.SimulationGrid(0)
    try:
        g.configure(scene.grid_config())
        s, e = scene.flat_materials()
        g.upload_materials(s, e, scene.subcell)
        g.set_sources(scene.sources)
        g.set_kernel(kernel)
        if absorber == "cpml":
            g.set_absorber(E.ABSORBER_CPML)  # default profile over the same 10-cell slabs
        st = {"f32": E.STORAGE_F32, "f16": E.STORAGE_F16, "bf16": E.STORAGE_BF16}[storage]
        if st != E.STORAGE_F32:
            # E near the feeds peaks at O(1); B at O(1/8): stored values stay far below 65504
            g.set_storage(st, [0, 0, 0, -3, -3, -3])
        ff = None
        if monitors and name != "vacuum":
            if scene.ports:
                pm = E.PortMonitor(g)
                pm.configure(scene.ports, scene.far_field.averaging_periods, scene.sweep, scene.cell_size)
            ff = E.FarFieldMonitor(g)
            ff.configure(scene.far_field)
            wl = scene.far_field.wavelength
            ff.set_wavelengths([wl, wl * 0.75, wl * 1.5])
        nx, ny, nz = g.size
        cells = nx * ny * nz
        g.step(warmup)
        g.synchronize()
        # bracket the Yee launches of at least MIN_TIMED_LAUNCHES timed steps (an event pair costs
        # the stream a few microseconds: every step at large sizes, every k-th step at small ones)
        every = 1 if cells >= 256 ** 3 * 8 else max(1, steps // (2 * MIN_TIMED_LAUNCHES))
        every = max(1, min(every, steps // MIN_TIMED_LAUNCHES)) if steps >= MIN_TIMED_LAUNCHES else 1
        g.kernel_timing(every)
        t0 = time.perf_counter()
        g.step(steps)
        g.synchronize()
        wall = time.perf_counter() - t0
        yee_ms, launches = g.kernel_timing(0)
        samples = g.kernel_timing_samples()
        Eg, _ = g.fields(E.LAYOUT_SOA)
        finite = bool(np.all(np.isfinite(Eg)))
        del Eg
        bpc = BYTES_PER_CELL if st == E.STORAGE_F32 else BYTES_PER_CELL // 2
        med_ms = statistics.median(samples) if samples else yee_ms / max(launches, 1)
        aux = 0
        if absorber == "cpml":
            slab = 2 * scene.pml + 1
            aux = CPML_AUX_BYTES * slab * (ny * nz + nx * nz + nx * ny)
        achieved = (bpc * cells + aux) / (med_ms * 1e-3) / 1e9
        key = f"{name}_{n}_{storage}_k{kernel}" + ("_cpml" if absorber == "cpml" else "")
        res = {
            "value": round(cells * steps / wall / 1e6, 2),
            "ms_per_step": round(wall / steps * 1e3, 5),
            "workload": workload_text(name, n, storage, monitors)
                        + (", convolutional PML in the absorber slabs" if absorber == "cpml" else ""),
            "grid": [nx, ny, nz], "courant": round(g.dt, 6), "material_classes": g.material_classes(),
            "kernel": kernel, "field_storage": storage, "absorber": absorber,
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_for(key),
                         "kernel": g.yee_kernel_name(), "median_launch_ms": round(med_ms, 5),
                         "mean_launch_ms": round(yee_ms / max(launches, 1), 5),
                         "min_launch_ms": round(min(samples), 5) if samples else None,
                         "max_launch_ms": round(max(samples), 5) if samples else None,
                         "launches": launches, "timed_every": every, "bytes_per_cell": bpc,
                         **({"cpml_aux_bytes": aux} if aux else {}),
                         "traffic_key": key},
        }
        if pattern and ff is not None:
            # far-field pattern of the workload: a 181 x 361 (theta, phi) grid at the three wavelengths
            dirs = E.FarFieldMonitor.theta_phi_grid(181, 361)
            ff.radiation(dirs[:256], 0)  # warm-up
            t_p = time.perf_counter()
            for f in range(3):
                ff.radiation(dirs, f)
            dt_p = time.perf_counter() - t_p
            pairs = 3 * dirs.shape[0] * ff.n_points
            res["ntff_pattern"] = {"directions": int(dirs.shape[0]), "points": int(ff.n_points), "wavelengths": 3,
                                   "seconds": round(dt_p, 4), "pairs_per_s": round(pairs / dt_p, 1),
                                   "flop_per_pair": "sincos + 12 FMA in the direct-sum form",
                                   "method": "per-face phase-matrix products on v_mfma_f32_32x32x16_f16 "
                                             "(fp16 hi/lo operand pairs, FP32 accumulation)"}
        return res, finite
    finally:
        g.close()


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=768)
    ap.add_argument("--kernel", type=int, default=0, help="EMFD_KERNEL_* variant (| options << 8)")
    ap.add_argument("--no-monitors", action="store_true")
    ap.add_argument("--scene", choices=["array", "spheres", "vacuum"], default="array",
                    help="array: seeded spheres around a 4x4 thin-wire dipole array (the precision-form "
                         "scene, headline); spheres: spheres around one thin-wire dipole; vacuum: no materials")
    ap.add_argument("--storage", choices=["f32", "f16", "bf16"], default="f32",
                    help="field storage: FP32 planes, or 16-bit planes with FP32 updates")
    ap.add_argument("--absorber", choices=["graded", "cpml"], default="graded",
                    help="graded: the reference's conductivity absorber; cpml: the convolutional PML")
    ap.add_argument("--no-extras", action="store_true", help="skip the vacuum / CPML / 16-bit / 256^3 extra workloads")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the learned screening measurement (designs/s against one-at-a-time FDTD)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="only time a plain float4 device copy of MB megabytes (roofline calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the single-GPU figure is the per-GPU share; the sharding design is written up, not built)",
              file=sys.stderr)
        return 2

    from em_lab_amd import emfdtd as E

    if args.copy_probe:
        g = E.SimulationGrid(0)
        nbytes = args.copy_probe * 1000 * 1000
        ms = g.copy_probe(nbytes, max(args.steps, 1))
        res = {"copy_probe_bytes": nbytes, "avg_ms": round(ms, 5),
               "read_plus_write_GBs": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}
        for planes in (1, 2, 6):
            sms, moved = g.stream_probe(nbytes, planes, max(args.steps, 1))
            res[f"stream_{planes}in_{planes}out_GBs"] = round(moved / (sms * 1e-3) / 1e9, 1)
        print(json.dumps(res), flush=True)
        g.close()
        return 0

    head, ok = run_grid(args.scene, args.size, args.storage, args.kernel, args.steps, args.warmup,
                        not args.no_monitors, pattern=not args.no_monitors, absorber=args.absorber)
    if not ok:
        print("bench.py: non-finite fields", file=sys.stderr)
        return 2
    print(f"bench.py: headline {head['value']} Mcell/s, Yee frac {head['roofline']['frac']}", file=sys.stderr,
          flush=True)
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage == "f32" else f"{args.storage} storage, f32 update",
        "data": "synthetic",
        "config": {k: head[k] for k in ("workload", "grid", "courant", "material_classes", "kernel",
                                        "field_storage", "absorber")},
        "roofline": head["roofline"],
        "cpu_baseline": None,
    }
    if "ntff_pattern" in head:
        out["ntff_pattern"] = head["ntff_pattern"]
    # the GPU's own streaming ceiling for the Yee pass's access pattern (6 planes read, 6 written
    # at one index, no stencil), against which the achieved figure can also be read
    g = E.SimulationGrid(0)
    try:
        plane = args.size ** 3 * 4
        ms, moved = g.stream_probe(plane, 6, 5)
        stream = moved / (ms * 1e-3) / 1e9
        out["roofline"]["stream_probe"] = {"GBs": round(stream, 1), "frac_of_peak": round(stream / HBM_PEAK_GBS, 4),
                                           "pattern": "6 planes read + 6 written per index, float4, no stencil",
                                           "yee_achieved_over_stream": round(head["roofline"]["achieved"] / stream, 4)}
    finally:
        g.close()
    if not args.no_extras:
        extras = {}
        n_ex = max(8, min(args.steps, 40))
        for key, (scene, n, storage, mon, absorber) in {
                "vacuum_768": ("vacuum", args.size, "f32", False, "graded"),
                "cpml_768": (args.scene, args.size, "f32", True, "cpml"),
                "f16_768": (args.scene, args.size, "f16", True, "graded"),
                "spheres_256": ("spheres", 256, "f32", True, "graded")}.items():
            r, ok = run_grid(scene, n, storage, args.kernel, n_ex if n > 256 else 4 * n_ex, 5, mon, absorber=absorber)
            if not ok:
                print(f"bench.py: non-finite fields in {key}", file=sys.stderr)
                return 2
            extras[key] = {"value": r["value"], "ms_per_step": r["ms_per_step"], "workload": r["workload"],
                           "roofline": r["roofline"]}
            print(f"bench.py: {key} {r['value']} Mcell/s, Yee frac {r['roofline']['frac']}", file=sys.stderr,
                  flush=True)
        out["extras"] = extras
    if not args.no_surrogate and args.storage == "f32":
        from em_lab_amd import surrogate as S
        out["surrogate"] = S.screening_run(65536, 256, plain_designs=4)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.scene if args.scene != "vacuum" else "spheres")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
