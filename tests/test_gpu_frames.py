"""The headless `--frames` driver on the GPU: a Hertzian scene run to completion prints the
reference's report (a header every 100 frames and at the end, the far-field summary with its three
cuts, the timing line) and exits 0; non-finite fields, loaded from the host before the run, give
the reference's error line and exit status 2 (src/Instruments.cpp:271-275, src/Main.cpp:1157)."""
import math

import numpy as np
import pytest

from em_lab_amd import emfdtd as E
from em_lab_amd import frames as F

pytestmark = pytest.mark.gpu


def test_frames_run_reports_and_exits_zero():
    print("case: frames hertzian 48^3, 120 frames", flush=True)
    lines = []
    rc = F.run(["--scene", "hertzian", "--n", "48", "--frames", "120", "--probe", "centre:24,24,30"],
               out=lines.append)
    assert rc == 0
    headers = [ln for ln in lines if ln.startswith("[t=")]
    assert len(headers) == 2  # frame 100 and the end of the run
    assert headers[-1] == "[t=%8.1f step %7d]" % (240 * F.build_scene("hertzian", 48).courant, 240)
    probe = [ln for ln in lines if ln.startswith("  > probe centre")]
    assert len(probe) == 2 and "nan" not in probe[-1]
    far = [ln for ln in lines if ln.startswith("    far field [lambda=16.0 cells")]
    assert len(far) == 2
    d_max = float(far[-1].split("D_max=")[1].split(" dBi")[0])
    assert 0.5 < d_max < 3.0  # a short z dipole: about 1.76 dBi once settled
    assert sum(1 for ln in lines if "-cut: peak" in ln) == 6
    assert lines[-1].startswith("GPU timings: sim ") and lines[-1].endswith("ms/frame (2 steps), render 0.000 ms")


def test_frames_non_finite_fields_exit_2(capsys):
    print("case: frames hertzian 48^3 with a NaN field texel", flush=True)

    def poison(grid):
        e, _ = grid.fields(E.LAYOUT_SOA)
        nx, ny, nz = grid.size
        e[:, (24 * ny + 24) * nx + 24] = math.nan
        grid.set_fields(e, None, E.LAYOUT_SOA)

    lines = []
    rc = F.run(["--scene", "hertzian", "--n", "48", "--frames", "4", "--probe", "c:24,24,24"], after_setup=poison,
               out=lines.append)
    assert rc == 2
    assert "ERROR: non-finite field values detected at step 2" in capsys.readouterr().err
    assert lines[-1].startswith("GPU timings: sim ")


def test_frames_fed_dipole_sweep_and_flux():
    """The fed-dipole scene (ports with a sweep, a flux box, the far field) at 64^3: the sweep
    summary replaces the continuous-wave port lines, the flux line appears, and the run ends with
    status 0."""
    print("case: frames spheres_dipole 64^3, 100 frames", flush=True)
    lines = []
    rc = F.run(["--scene", "spheres_dipole", "--n", "64", "--frames", "100"], out=lines.append)
    assert rc == 0
    assert any(ln.startswith("    flux  ") for ln in lines)
    assert any(ln.startswith("    sweep: ") and "steps recorded" in ln for ln in lines)
    assert any(ln.startswith("    sweep dipole") and "best match" in ln for ln in lines)
    assert not any(ln.startswith("    port  ") for ln in lines)  # sweeping: no continuous-wave lines



def test_frames_coax_line():
    """The reference's coax validation scene headless (8 steps per frame, as its view block):
    the cable lines appear with the measured wave impedance and the matched load's reflection."""
    print("case: frames coax_line 232x64x64, 400 frames of 8 steps", flush=True)
    lines = []
    rc = F.run(["--scene", "coax_line", "--frames", "400", "--steps", "8"], out=lines.append)
    assert rc == 0
    head = [ln for ln in lines if ln.startswith("    cable: Z0 50.0 ohm (eps_r 2.46")]
    dt = float(np.float32(F.build_scene("coax_line").courant))
    assert len(head) == 5 and head[-1].endswith("load end, %.0f periods" % (3200 * dt / 48.0))  # 4 + the end
    waves = [ln for ln in lines if ln.startswith("      waves: Z0 measured ")]
    assert waves and float(waves[-1].split("Z0 measured ")[1].split(" ohm")[0]) == pytest.approx(50.0, abs=1.0)
    assert float(waves[-1].split("end Gamma ")[1].split(" @")[0]) < 0.08


def test_frames_radar_scene_reports_detection():
    """--scene radar: the calibration pass on the scene without its target, the target swapped in,
    the measurement pass, and printRadar's lines (src/Instruments.cpp:412-442) with the detection
    against the target's front."""
    print("case: frames radar, 2x4 dipole array, sphere", flush=True)
    lines = []
    rc = F.run(["--scene", "radar", "--frames", "600", "--steps", "8"], out=lines.append)
    assert rc == 0
    summary = [ln for ln in lines if ln.startswith("    radar: strongest echo")]
    assert summary, lines[-12:]
    det = [ln for ln in lines if ln.startswith("      #1 range")]
    assert det and "| sphere: front at" in det[-1]
    rng = float(det[-1].split("range")[1].split("mm")[0])
    front = float(det[-1].split("front at")[1].split("mm")[0])
    assert abs(rng - front) < 2.0  # mm (1 mm cells)


def test_frames_crpa_scene_reports_adaptation():
    print("case: frames crpa ideal, 2x2 crossed dipoles", flush=True)
    lines = []
    rc = F.run(["--scene", "crpa:ideal", "--frames", "400", "--steps", "8"], out=lines.append)
    assert rc == 0
    sinr = [ln for ln in lines if ln.startswith("    crpa (MVDR weights): SINR")]
    assert sinr, lines[-12:]
    applied = float(sinr[-1].split("applied")[1].split("dB")[0])
    steering = float(sinr[-1].split("steering")[1].split("dB")[0])
    assert applied > steering + 20.0
    assert sum(1 for ln in lines if ln.startswith("      jam ")) >= 2
