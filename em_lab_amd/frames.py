"""Headless run of a scene for a number of frames: EM_Lab's `--frames` mode over this engine.

Follows `Application::run` (src/Main.cpp:1043-1158) without the window: each frame advances
`--steps` steps (default 2, clamped to 0..64 as src/Main.cpp:102), reads the probes and flux
meters every second frame (`Instruments::measure`, src/Instruments.cpp:225-276), evaluates the
far field, the ports and the sweep every 50th frame (frame % 50 == 49), prints the report every
100 frames and, after the last frame, measures, evaluates and prints it once more
(`Instruments::printReport`, src/Instruments.cpp:443-493, with `printPorts` :288-363) and a
timing line. Exit status: 3 when the scene cannot be built (a script error in the reference),
2 when a probe or flux reading was not finite (src/Instruments.cpp:271-275), else 0. The
interactive one-second far-field timer of the reference is left out (the frame rule alone
decides, so a run is reproducible).

Scenes come from the Python scene builders (em_lab_amd.scene), since there is no Lua runtime
here: `hertzian` (64^3 Hertzian dipole, far field), `spheres_dipole` (256^3 seeded spheres and
a fed thin-wire dipole with sweep, flux box and far field), `spheres_array` (the 4x4 array
scene), `array:N` (design N of the batched generator).

    python -m em_lab_amd.frames --scene hertzian --frames 200 [--steps 2] [--n 64]
        [--probe NAME:X,Y,Z ...] [--sweep-csv FILE] [--pattern-csv FILE]
"""
from __future__ import annotations

import argparse
import math
import sys
import time
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

ETA0 = 376.730313668  # src/Instruments.cpp kEta0
C0 = 299792458.0
NO_INSTRUMENTS_HINT = (
    "No instruments in this scene, so there are no readings to report. Add probe{}, flux_plane{}, "
    "flux_box{} or far_field{} to the script, or feed an antenna, to measure fields, power, impedance or "
    "radiation patterns.")


def build_scene(name: str, n: Optional[int] = None):
    """The named scene of the builders above; raises on an unknown name or a failing builder."""
    from em_lab_amd import scene as S
    if name == "hertzian":
        return S.hertzian_vacuum(n or 64, 8)
    if name == "spheres_dipole":
        return S.spheres_dipole(n=n or 256)
    if name == "spheres_array":
        return S.spheres_array(n=n or 192)
    if name.startswith("array:"):
        return S.draw_array_scene(int(name.split(":", 1)[1]))[1]
    if name == "radar" or name.startswith("radar:"):
        # a 2 x 4 dipole array as a pulsed radar (scenes/16_array_radar.lua's processing):
        # "radar:sphere" (default) or "radar:plate"
        return S.radar_dipole_array(target=name.split(":", 1)[1] if ":" in name else "sphere")
    if name in ("crpa", "crpa:ideal", "crpa:blind"):
        # a 2 x 2 crossed-dipole CRPA with a satellite and two jammers (scenes/17_crpa_gps.lua's
        # processing), blind (default) or ideal adaptation
        return S.crpa_crossed_dipoles(blind=not name.endswith(":ideal"))
    if name == "coax_line" or name.startswith("coax_line:"):
        # scenes/validation/coax_line.lua: a 150-cell 50 ohm line, lambda 48, terminated by a
        # matched load, an open, a short or 100 ohm ("coax_line:open", "coax_line:100")
        term = name.split(":", 1)[1] if ":" in name else "load"
        if term not in ("load", "open", "short", "100"):
            raise ValueError(f"unknown coax_line termination '{term}'")
        sc = S.Scene(size=(232, 64, 64), pml=16, wavelength=48.0)
        sc.coax((36, 32, 32), axis=0, sign=1, length=150, z0=50.0, termination="load" if term == "100" else term,
                load=100.0 if term == "100" else None)
        return sc
    raise ValueError(f"unknown scene '{name}'")


@dataclass
class ProbeState:
    name: str
    position: tuple
    sample: Optional[dict] = None


@dataclass
class FluxState:
    surface: object
    instantaneous: float = 0.0
    time_averaged: float = 0.0
    peak_abs: float = 0.0
    last_time: float = -1.0


def update_flux(f: FluxState, P: float, t: float, averaging_time: float) -> None:
    """One flux reading at time t: instantaneous value, running mean with time constant
    `averaging_time` (restarted when time runs backwards) and peak magnitude
    (src/Instruments.cpp:256-266)."""
    if f.last_time < 0.0 or t < f.last_time:
        f.peak_abs = 0.0
        f.time_averaged = P
    else:
        a = 1.0 - math.exp(-(t - f.last_time) / averaging_time)
        f.time_averaged += a * (P - f.time_averaged)
    f.instantaneous = P
    f.peak_abs = max(f.peak_abs, abs(P))
    f.last_time = t


@dataclass
class Instruments:
    """The measuring side of a headless run: host mirror of src/Instruments.cpp's state."""
    grid: object
    probes: list = field(default_factory=list)
    fluxes: list = field(default_factory=list)
    ports: object = None       # PortMonitor or None
    port_setups: list = field(default_factory=list)
    far_field: object = None   # FarFieldMonitor or None
    ff_setup: object = None
    cell_size: float = 1.0e-3
    averaging_time: float = 6.0
    unstable: bool = False
    measure_time: float = 0.0
    measure_step: int = 0
    readings: list = field(default_factory=list)
    sweep: Optional[dict] = None
    pattern: Optional[dict] = None
    start_time: float = 0.0
    err: Callable[[str], None] = lambda s: print(s, file=sys.stderr, flush=True)

    coax: object = None
    radar: object = None
    crpa: object = None

    def empty(self) -> bool:
        return (not self.probes and not self.fluxes and self.far_field is None and self.ports is None
                and self.coax is None and self.radar is None and self.crpa is None)

    # src/Instruments.cpp:393-410
    def coax_lines(self) -> list:
        out = []
        if self.coax is None or not self.coax.enabled():
            return out
        periods = self.coax.periods()
        for c, r in zip(self.coax.lines, self.coax.polled):
            out.append("    %s: Z0 %.1f ohm (eps_r %.2f, velocity factor %.2f), %s end, %.0f periods"
                       % (c.name, c.z0_ohm, c.epsilon_r, 1.0 / math.sqrt(c.epsilon_r), c.termination, periods))
            if not r.valid:
                continue
            g = r.gamma_end
            out.append("      waves: Z0 measured %.1f ohm, end Gamma %.3f @ %+.0f deg (RL %.1f dB), VSWR %.2f, "
                       "Z_end %.1f %+.1fj ohm" % (r.z0_measured, abs(g), math.degrees(math.atan2(g.imag, g.real)),
                                                 r.return_loss_db, r.vswr, r.z_end.real, r.z_end.imag))
            out.append("      forward %.0f%% of available, delivered %.0f%% of forward; shield outside current %.1f dB "
                       "vs line current; fit residual %.1e" % (100.0 * r.forward_fraction, 100.0 * (1.0 - abs(g) ** 2),
                                                              max(r.common_mode_db, -200.0), r.fit_residual))
        return out

    def needs_every_step(self) -> bool:
        return self.far_field is not None or self.ports is not None

    def sweeping(self) -> bool:
        return self.ports is not None and self.ports._sweep is not None and bool(self.port_setups)

    # src/Instruments.cpp:225-276
    def measure(self) -> None:
        from em_lab_amd import emfdtd as E
        g = self.grid
        if not self.probes and not self.fluxes:
            self.measure_time, self.measure_step = g.time, g.step_count
            return
        an = E.FieldAnalyzer(g)
        self.measure_time, self.measure_step = g.time, g.step_count
        finite = True
        if self.probes:
            samples = an.probe([p.position for p in self.probes])
            for p, s in zip(self.probes, samples):
                p.sample = s
                finite = finite and math.isfinite(float(np.dot(s["E"], s["E"]))) and \
                    math.isfinite(float(np.dot(s["B"], s["B"])))
        if self.fluxes:
            P = an.measure_flux([f.surface for f in self.fluxes])
            for f, v in zip(self.fluxes, P):
                v = float(v)
                finite = finite and math.isfinite(v)
                update_flux(f, v, self.measure_time, self.averaging_time)
        if not finite and not self.unstable:
            self.unstable = True
            self.err(f"ERROR: non-finite field values detected at step {self.measure_step}")

    # src/Instruments.cpp:278-281 and the sweep (src/Ports.cpp:187-278)
    def evaluate_far_field(self) -> None:
        if self.far_field is not None:
            self.pattern = self.far_field.evaluate()
        if self.ports is not None:
            self.readings = self.ports.evaluate()

    def evaluate_sweep(self) -> None:
        from em_lab_amd import emfdtd as E
        try:
            self.sweep = self.ports.evaluate_sweep(self.grid.dt)
        except E.EmfdError:
            pass  # fewer than 16 recorded steps: the previous result stands (src/Ports.cpp:188)

    # src/Instruments.cpp:443-493
    def report_lines(self) -> list:
        g = self.grid
        if self.empty():
            return ["[t=%8.1f step %7d] %s" % (g.time, g.step_count, NO_INSTRUMENTS_HINT)]
        out = ["[t=%8.1f step %7d]" % (self.measure_time, self.measure_step)]
        for i, p in enumerate(self.probes):
            s = p.sample or dict(E=np.zeros(3), B=np.zeros(3), S=np.zeros(3), sigma=0.0, epsilon_r=1.0)
            E_, B_, S_ = (np.asarray(s[k], np.float32) for k in ("E", "B", "S"))
            em = float(np.sqrt(np.float32(np.dot(E_, E_))))
            bm = float(np.sqrt(np.float32(np.dot(B_, B_))))
            out.append("  %c probe %-14s (%6.1f %6.1f %6.1f): |E|=%.5f E=(%+.4f %+.4f %+.4f) |B|=%.5f "
                       "S=(%+.2e %+.2e %+.2e) sigma=%.3g eps_r=%.2f"
                       % ('>' if i == 0 else ' ', p.name, p.position[0], p.position[1], p.position[2], em,
                          E_[0], E_[1], E_[2], bm, S_[0], S_[1], S_[2], s["sigma"], s["epsilon_r"]))
        for f in self.fluxes:
            out.append("    flux  %-28s P=%+.4e  <P>=%+.4e  peak|P|=%.4e"
                       % (f.surface.label, f.instantaneous, f.time_averaged, f.peak_abs))
        out += self.port_lines(4)
        if self.radar is not None:
            from em_lab_amd.radar import report_lines as radar_lines
            out += radar_lines(self.radar)
        if self.crpa is not None:
            out += self.crpa.report_lines()
        out += self.coax_lines()
        if self.far_field is not None:
            st = self.ff_setup
            settle = max(g.time - self.start_time, 0.0) / max(st.wavelength, 1.0)
            pat = self.pattern
            if not pat or not pat.get("valid"):
                out.append("    far field: accumulating (%.1f periods)..." % settle)
            else:
                d = np.asarray(pat["max_direction"], np.float32)
                theta = math.degrees(math.acos(min(max(float(d[2]), -1.0), 1.0)))
                phi = math.degrees(math.atan2(float(d[1]), float(d[0])))
                still = ", still settling" if settle < 3.0 * st.averaging_periods else ""
                out.append("    far field [lambda=%.1f cells, %.0f periods averaged%s]: P_rad=%.4e  D_max=%.2f dBi "
                           "(%.2f) @ theta=%.0f phi=%.0f"
                           % (st.wavelength, settle, still, pat["radiated_power"], pat["max_directivity_dbi"],
                              10.0 ** (pat["max_directivity_dbi"] / 10.0), theta, phi))
                for name in ("xy", "xz", "yz"):
                    c = pat["cuts"][name]
                    bw = "omni" if c["beamwidth_deg"] >= 359.0 else "%.1f deg" % c["beamwidth_deg"]
                    out.append("      %s-cut: peak %6.2f dBi @ %5.1f deg, HPBW %s, front/back %.1f dB"
                               % (name, c["peak_dbi"], c["peak_angle_deg"], bw, c["front_to_back_db"]))
        return out

    # src/Instruments.cpp:288-363
    def port_lines(self, max_lines: int) -> list:
        if self.ports is None:
            return []
        out = []
        ports, rd = self.port_setups, self.readings
        periods = max(self.grid.time - self.start_time, 0.0) / ports[0].wavelength if ports else 0.0
        accepted = sum(r.accepted_power for r in rd)
        available = sum(r.available_power for r in rd)
        cw_lines = 0 if self.sweeping() else max_lines
        if periods < 12.0 and cw_lines > 0:
            out.append("    ports: averaging (%.0f of ~12 periods), impedances still settling" % periods)
        for i, p in enumerate(ports[:cw_lines]):
            if i >= len(rd) or not rd[i].valid:
                out.append("    port  %-14s (no current yet)" % p.name)
                continue
            r = rd[i]
            if p.voltage_feed:
                out.append("    port  %-14s Z = %7.1f %+7.1fj ohm | Gamma %.3f @ %+5.0f deg (S11 %6.1f dB), VSWR %5.2f vs "
                           "%.0f ohm | P_acc %.3e of %.3e avail (%.0f%%)"
                           % (p.name, r.z_ohm.real, r.z_ohm.imag, abs(r.gamma),
                              math.atan2(r.gamma.imag, r.gamma.real) * 57.29577951308232, -r.return_loss_db, r.vswr,
                              p.impedance_ohm, r.accepted_power, r.available_power,
                              100.0 * r.accepted_power / r.available_power if r.available_power > 0.0 else 0.0))
            else:
                out.append("    port  %-14s Z = %7.1f %+7.1fj ohm | current feed %.3g | P_acc %.3e"
                           % (p.name, r.z_ohm.real, r.z_ohm.imag, p.drive, r.accepted_power))
        if cw_lines > 0 and len(ports) > max_lines:
            out.append("    ... %d more ports (press P for all) | total P_acc %.3e of %.3e available"
                       % (len(ports) - max_lines, accepted, available))
        if self.sweeping():
            sw = self.sweep

            def ghz(wl):
                return C0 / (wl * self.cell_size) * 1e-9

            if not sw:
                out.append("    sweep: recording the pulse response...")
            else:
                res = sw["residual"]
                tail = (" (buffer full: results approximate)" if sw["full"] else " (still ringing: results settling)") \
                    if res > 1e-3 else " (settled)"
                out.append("    sweep: %d steps recorded, ring-down %.1e of peak%s" % (sw["samples"], res, tail))
                for i, ps in enumerate(sw["ports"][:max_lines]):
                    b = ps["points"][ps["best"]]
                    line = ("    sweep %-12s best match %.4f GHz (lambda %.2f cells): S11 %.1f dB, VSWR %.2f, "
                            "Z = %.1f %+.1fj ohm" % (ports[i].name, ghz(ps["best_wavelength"]), ps["best_wavelength"],
                                                     b[3], b[4], b[1], b[2]))
                    if ps["has_band"]:
                        f0, fl, fh = ghz(ps["best_wavelength"]), ghz(ps["band_low_wavelength"]), \
                            ghz(ps["band_high_wavelength"])
                        line += " | -10 dB band %.4f-%.4f GHz (%.1f%%)" % (fl, fh, 100.0 * (fh - fl) / f0)
                    else:
                        line += " | never below -10 dB in the sweep"
                    out.append(line)
        pat = self.pattern
        if pat and pat.get("valid") and accepted > 0.0:
            prad = pat["radiated_power"]
            if len(ports) == 1 and rd and rd[0].valid:
                i2 = abs(rd[0].I) ** 2
                out.append("    radiation resistance R_rad = 2 P_rad / |I|^2 = %.1f ohm, radiation efficiency "
                           "P_rad / P_acc = %.0f%%" % (2.0 * prad / i2 * ETA0, 100.0 * prad / accepted))
            else:
                out.append("    radiation efficiency P_rad / total P_acc = %.0f%%" % (100.0 * prad / accepted))
        return out


def parse_probe(text: str):
    name, _, xyz = text.rpartition(":")
    pos = tuple(float(v) for v in xyz.split(","))
    if len(pos) != 3:
        raise ValueError(f"--probe expects NAME:X,Y,Z (got '{text}')")
    return name or "probe", pos


def run(argv=None, after_setup: Optional[Callable] = None, out=print) -> int:
    """The headless run; returns the exit status. `after_setup(grid)` runs once the grid and the
    instruments are configured (a host-side hook, e.g. to load fields)."""
    ap = argparse.ArgumentParser(prog="python -m em_lab_amd.frames", description=__doc__.split("\n\n")[0])
    ap.add_argument("--scene", default="hertzian")
    ap.add_argument("--n", type=int, default=None, help="lattice size for the cubic scenes")
    ap.add_argument("--frames", type=int, default=100)
    ap.add_argument("--steps", type=int, default=2, help="steps per frame (0..64)")
    ap.add_argument("--probe", action="append", default=[], metavar="NAME:X,Y,Z")
    ap.add_argument("--sweep-csv", default=None)
    ap.add_argument("--pattern-csv", default=None)
    args = ap.parse_args(argv)
    steps = min(max(args.steps, 0), 64)
    try:
        sc = build_scene(args.scene, args.n)
        probes = [parse_probe(p) for p in args.probe]
    except Exception as e:  # the reference's script error
        print(f"ERROR: scene '{args.scene}': {e}", file=sys.stderr, flush=True)
        return 3

    from em_lab_amd import emfdtd as E
    g = E.SimulationGrid(0)
    try:
        g.configure(sc.grid_config())
        s, e = sc.flat_materials()
        radar = getattr(sc, "radar", None)
        crpa = getattr(sc, "crpa", None)
        if radar is not None:
            # the calibration pass runs on the scene without its targets (src/Main.cpp:472-505)
            bs, be = sc.background_materials()
            g.upload_materials(bs, be, sc.subcell)
        else:
            g.upload_materials(s, e, sc.subcell)
        g.set_sources(sc.sources)
        size = g.size
        inst = Instruments(g, cell_size=sc.cell_size, averaging_time=3.0 * max(float(sc.wavelength), 2.0))
        for name, pos in probes:
            inst.probes.append(ProbeState(name, tuple(min(max(v, 0.0), size[k] - 1.0) for k, v in enumerate(pos))))
        inst.fluxes = [FluxState(f) for f in (sc.fluxes or [])]
        if radar is not None or crpa is not None:
            # one port monitor per context: the radar's / CRPA's recorder takes it
            if radar is not None:
                from em_lab_amd.radar import RadarMonitor
                inst.radar = RadarMonitor(g)
                inst.radar.configure(radar, sc.ports, sc.sources, sc.ports[0].wavelength if sc.ports else sc.wavelength,
                                     sc.cell_size, sc.targets)
            else:
                from em_lab_amd.crpa import CrpaMonitor
                inst.crpa = CrpaMonitor(g)
                inst.crpa.configure(crpa, sc.emitters, sc.sources, sc.ports,
                                    sc.ports[0].wavelength if sc.ports else sc.wavelength, sc.cell_size)
        elif sc.ports:
            inst.ports = E.PortMonitor(g)
            avg = sc.far_field.averaging_periods if sc.far_field else 4.0
            inst.ports.configure(sc.ports, avg, sc.sweep, sc.cell_size)
            inst.port_setups = list(sc.ports)
        if sc.far_field:
            inst.far_field = E.FarFieldMonitor(g)
            inst.ff_setup = inst.far_field.configure(sc.far_field)
        if getattr(sc, "coaxes", None):
            from em_lab_amd.coax import CoaxMonitor
            inst.coax = CoaxMonitor(g)
            inst.coax.configure(sc.coaxes)
        if after_setup:
            after_setup(g)
        if inst.empty():
            out(NO_INSTRUMENTS_HINT)
        sim_ms = 0.0
        for frame in range(max(args.frames, 0)):
            if inst.crpa is not None:
                inst.crpa.update()  # switches the transmitting sources between passes (src/Main.cpp:1067)
            if steps > 0:
                t = time.perf_counter()
                g.step(steps)
                g.synchronize()
                sim_ms = (time.perf_counter() - t) * 1e3
            if inst.radar is not None:
                # swap in the targets once the calibration pass is recorded (src/Main.cpp:1079-1087)
                inst.radar.update()
                if inst.radar.calibration_complete():
                    g.upload_materials(s, e, sc.subcell)
                    g.reset_fields()
                    inst.radar.begin_measurement()
            if frame % 2 == 0 or steps == 0:
                inst.measure()
            if inst.coax is not None:
                inst.coax.poll()  # once per frame, as advanceFarFieldEvaluation (src/Instruments.hpp:70-74)
            if inst.needs_every_step() and frame % 50 == 49:
                inst.evaluate_far_field()
                if inst.sweeping():
                    inst.evaluate_sweep()
            done = frame + 1
            if done % 100 == 0 and not inst.empty():
                for line in inst.report_lines():
                    out(line)
        inst.measure()
        inst.evaluate_far_field()
        if inst.sweeping():
            inst.evaluate_sweep()
        if not inst.empty():
            for line in inst.report_lines():
                out(line)
        out("GPU timings: sim %.3f ms/frame (%d steps), render %.3f ms" % (sim_ms, steps, 0.0))
        if args.sweep_csv and inst.sweep:
            E.write_sweep_csv(args.sweep_csv, [p.name for p in inst.port_setups], inst.sweep, sc.cell_size)
        if args.pattern_csv and inst.pattern:
            E.write_pattern_csv(args.pattern_csv, inst.pattern)
        return 2 if inst.unstable else 0
    finally:
        g.close()


def main(argv=None) -> int:
    return run(argv)


if __name__ == "__main__":
    sys.exit(main())
