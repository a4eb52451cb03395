"""Host side of the headless `--frames` driver (em_lab_amd.frames): the report lines against the
reference's printf formats (src/Instruments.cpp:288-363,443-493) on hand-chosen readings, the flux
running mean (src/Instruments.cpp:256-266), probe parsing and the scene-error exit status."""
import math
from types import SimpleNamespace

import numpy as np
import pytest

from em_lab_amd import frames as F
from em_lab_amd.emfdtd import FarFieldSetup, FluxSurface, PortReading, PortSetup, SweepSetup


def _instruments(sweep=None):
    grid = SimpleNamespace(time=123.45, step_count=216, dt=0.57)
    ports = SimpleNamespace(_sweep=sweep)
    inst = F.Instruments(grid, cell_size=1.0e-3)
    inst.measure_time, inst.measure_step = 123.45, 216
    inst.probes = [F.ProbeState("p1", (10.0, 20.0, 30.0),
                                dict(E=np.array([0.1, -0.2, 0.05], np.float32), B=np.array([0.0, 0.001, 0.0], np.float32),
                                     S=np.array([1e-4, 0.0, -2.5e-5], np.float32), sigma=0.0, epsilon_r=1.0))]
    fx = F.FluxState(FluxSurface.box((8, 8, 8), (40, 40, 40), label="box"))
    fx.instantaneous, fx.time_averaged, fx.peak_abs = 1.5e-3, -2.25e-4, 3e-3
    inst.fluxes = [fx]
    inst.ports = ports
    inst.port_setups = [PortSetup(name="feed", wavelength=16.0)]
    inst.readings = [PortReading(True, 0j, 0.01 + 0j, 73.1 + 42.5j, 0.3 + 0.4j, 3.0, 6.0206, 1.25e-3, 2.5e-3)]
    inst.far_field = object()
    inst.ff_setup = FarFieldSetup(wavelength=16.0, averaging_periods=4.0)
    inst.pattern = dict(valid=True, radiated_power=5e-4, max_directivity_dbi=1.76,
                        max_direction=np.array([1.0, 0.0, 0.0], np.float32),
                        cuts=dict(xy=dict(peak_dbi=1.76, peak_angle_deg=0.0, beamwidth_deg=360.0, front_to_back_db=0.0),
                                  xz=dict(peak_dbi=1.70, peak_angle_deg=90.0, beamwidth_deg=90.0, front_to_back_db=0.0),
                                  yz=dict(peak_dbi=1.70, peak_angle_deg=90.0, beamwidth_deg=90.0,
                                          front_to_back_db=0.0)))
    return inst


def test_report_lines_continuous_wave():
    lines = _instruments().report_lines()
    assert lines == [
        "[t=   123.5 step     216]",
        "  > probe p1             (  10.0   20.0   30.0): |E|=0.22913 E=(+0.1000 -0.2000 +0.0500) |B|=0.00100 "
        "S=(+1.00e-04 +0.00e+00 -2.50e-05) sigma=0 eps_r=1.00",
        "    flux  box                          P=+1.5000e-03  <P>=-2.2500e-04  peak|P|=3.0000e-03",
        "    ports: averaging (8 of ~12 periods), impedances still settling",
        "    port  feed           Z =    73.1   +42.5j ohm | Gamma 0.500 @   +53 deg (S11   -6.0 dB), VSWR  3.00 "
        "vs 50 ohm | P_acc 1.250e-03 of 2.500e-03 avail (50%)",
        "    radiation resistance R_rad = 2 P_rad / |I|^2 = 3767.3 ohm, radiation efficiency P_rad / P_acc = 40%",
        "    far field [lambda=16.0 cells, 8 periods averaged, still settling]: P_rad=5.0000e-04  D_max=1.76 dBi "
        "(1.50) @ theta=90 phi=0",
        "      xy-cut: peak   1.76 dBi @   0.0 deg, HPBW omni, front/back 0.0 dB",
        "      xz-cut: peak   1.70 dBi @  90.0 deg, HPBW 90.0 deg, front/back 0.0 dB",
        "      yz-cut: peak   1.70 dBi @  90.0 deg, HPBW 90.0 deg, front/back 0.0 dB",
    ]


def test_report_lines_sweep():
    inst = _instruments(sweep=SweepSetup(12.0, 24.0, 3))
    inst.probes, inst.fluxes = [], []
    inst.pattern = None
    points = np.array([[24.0, 30.0, 10.0, -5.0, 3.5], [20.0, 48.0, -3.5, -25.3, 1.12], [12.0, 90.0, 5.0, -3.0, 5.8]])
    inst.sweep = dict(residual=5e-4, samples=3000, full=False,
                      ports=[dict(points=points, best=1, best_wavelength=20.0, has_band=True,
                                  band_low_wavelength=21.0, band_high_wavelength=19.0)])
    lines = inst.report_lines()
    assert lines[1:] == [
        "    sweep: 3000 steps recorded, ring-down 5.0e-04 of peak (settled)",
        "    sweep feed         best match 14.9896 GHz (lambda 20.00 cells): S11 -25.3 dB, VSWR 1.12, Z = 48.0 -3.5j ohm"
        " | -10 dB band 14.2758-15.7786 GHz (10.0%)",
        "    far field: accumulating (7.7 periods)...",
    ]
    inst.sweep["ports"][0]["has_band"] = False
    inst.sweep["residual"], inst.sweep["full"] = 2e-2, True
    lines = inst.report_lines()
    assert lines[1] == "    sweep: 3000 steps recorded, ring-down 2.0e-02 of peak (buffer full: results approximate)"
    assert lines[2].endswith("Z = 48.0 -3.5j ohm | never below -10 dB in the sweep")
    inst.sweep = None
    assert inst.report_lines()[1] == "    sweep: recording the pulse response..."


def test_report_without_instruments():
    inst = F.Instruments(SimpleNamespace(time=5.0, step_count=9, dt=0.5))
    assert inst.empty() and not inst.needs_every_step()
    assert inst.report_lines() == ["[t=     5.0 step       9] " + F.NO_INSTRUMENTS_HINT]


def test_flux_running_mean():
    f = F.FluxState(None)
    F.update_flux(f, 2.0, 10.0, 6.0)  # first reading: the mean starts at it
    assert (f.instantaneous, f.time_averaged, f.peak_abs, f.last_time) == (2.0, 2.0, 2.0, 10.0)
    F.update_flux(f, -4.0, 16.0, 6.0)  # one time constant later: 1 - 1/e of the way
    assert f.time_averaged == pytest.approx(2.0 + (1.0 - math.exp(-1.0)) * -6.0)
    assert f.peak_abs == 4.0 and f.instantaneous == -4.0
    F.update_flux(f, 1.0, 3.0, 6.0)  # time ran backwards (a reset): start over
    assert (f.time_averaged, f.peak_abs) == (1.0, 1.0)


def test_probe_parsing_and_scene_errors():
    assert F.parse_probe("tip:1,2.5,3") == ("tip", (1.0, 2.5, 3.0))
    assert F.parse_probe("4,5,6") == ("probe", (4.0, 5.0, 6.0))
    with pytest.raises(ValueError):
        F.parse_probe("p:1,2")
    with pytest.raises(ValueError):
        F.build_scene("no-such-scene")
    # a scene that cannot be built ends the run with status 3 before any device work
    assert F.run(["--scene", "no-such-scene", "--frames", "1"]) == 3
    assert F.run(["--scene", "hertzian", "--probe", "bad", "--frames", "1"]) == 3


def test_coax_report_lines():
    """printCoax (src/Instruments.cpp:393-410) on a hand-chosen reading."""
    from em_lab_amd.coax import CoaxReading, CoaxSetup
    inst = F.Instruments(SimpleNamespace(time=480.0, step_count=842, dt=0.57))
    c = CoaxSetup(name="cable", z0_ohm=50.0, epsilon_r=2.25, termination="load", wavelength=48.0)
    r = CoaxReading(valid=True, z0_measured=49.96, gamma_end=complex(0.0, -0.0583), return_loss_db=24.69, vswr=1.1238,
                    z_end=complex(48.9, -5.7), forward_fraction=0.9948, common_mode_db=-250.0, fit_residual=1.6e-3)
    inst.coax = SimpleNamespace(enabled=lambda: True, periods=lambda: 10.0, lines=[c, c],
                                polled=[r, CoaxReading()])
    assert inst.coax_lines() == [
        "    cable: Z0 50.0 ohm (eps_r 2.25, velocity factor 0.67), load end, 10 periods",
        "      waves: Z0 measured 50.0 ohm, end Gamma 0.058 @ -90 deg (RL 24.7 dB), VSWR 1.12, Z_end 48.9 -5.7j ohm",
        "      forward 99% of available, delivered 100% of forward; shield outside current -200.0 dB vs line current; "
        "fit residual 1.6e-03",
        "    cable: Z0 50.0 ohm (eps_r 2.25, velocity factor 0.67), load end, 10 periods",
    ]
    with pytest.raises(ValueError):
        F.build_scene("coax_line:lossy")
