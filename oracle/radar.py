"""CPU restatement of the radar's imaging loop: test infrastructure only (the checker of
libemfdtd.so's emfd_radar_backproject, never part of the product path).

Follows RadarMonitor::computeImage's back-projection (src/Radar.cpp:622-658) point by point in
scalar Python: direction(az, el) (src/Radar.cpp:67-70) and positions in float32, the transmit
delay txDelay (src/Radar.cpp:86-104) and the coherent sum in float64, the echoes complex64
interpolated linearly between decimated samples. Sized for small images (a few hundred points).

Pins (tests/test_radar_cpu.py): direction(0, 0) is the boresight, direction(90, 0) the azimuth
axis and direction(0, 90) the elevation axis; one transmitter gives txDelay = |P - tx| whatever
its amplitude and off-boresight angle; one element with constant echoes c and weight w gives
power w^2 |c|^2 wherever its sample index is inside the record, 0 outside it and -1 outside the
simulated volume; two elements with echoes built as e^{-j w t} at their own delays to a chosen
point add coherently there (power (w1 + w2)^2).
"""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32


def direction(axes, az_deg: float, el_deg: float) -> np.ndarray:
    axes = np.asarray(axes, np.float32).reshape(3, 3)
    a = f32(f32(az_deg) * f32(0.017453292519943295))
    e = f32(f32(el_deg) * f32(0.017453292519943295))
    sa, ca, se, ce = np.sin(a), np.cos(a), np.sin(e), np.cos(e)
    return (f32(sa * ce) * axes[0] + se * axes[1]) + f32(ca * ce) * axes[2]


def _len(v) -> float:
    v = np.asarray(v, np.float32)
    return float(np.sqrt(f32(f32(v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])))


def tx_delay(P, tx, axes, tx_pos, tx_amp, wavelength: float) -> float:
    axes = np.asarray(axes, np.float32).reshape(3, 3)
    w = 2.0 * math.pi / wavelength
    a = 0j
    ad = 0j
    mag = 0.0
    for j in range(len(tx_amp)):
        v = (np.asarray(P, np.float32) - np.asarray(tx_pos[j], np.float32)).astype(np.float32)
        d = max(_len(v), 0.5)
        g = max(float(f32(f32(v[0] * axes[2][0] + v[1] * axes[2][1]) + v[2] * axes[2][2])) / d, 0.0)
        s = float(tx_amp[j]) * g / d
        term = s * complex(math.cos(-w * d), math.sin(-w * d))
        a += term
        ad += d * term
        mag += s
    if abs(a) < 1e-3 * mag:
        return _len(np.asarray(P, np.float32) - np.asarray(tx, np.float32))
    return (ad / a).real


def back_project(setup: dict, z, element_pos, weights, tx_pos, tx_amp) -> np.ndarray:
    """power [nr][nel][naz] as emfd_radar_backproject (see the module docstring)."""
    z = np.asarray(z, np.complex64)
    m, n_el = z.shape
    axes = np.asarray(setup["axes"], np.float32).reshape(3, 3)
    tx = np.asarray(setup["tx"], np.float32)
    lo = np.asarray(setup["interior_min"], np.float32)
    hi = np.asarray(setup["interior_max"], np.float32)
    w = 2.0 * math.pi / float(f32(setup["wavelength"]))
    dt, sdt = float(setup["dt"]), float(setup["sample_dt"])
    nr, nel, naz = int(setup["nr"]), int(setup["nel"]), int(setup["naz"])
    out = np.zeros((nr, nel, naz))
    for ie in range(nel):
        for ia in range(naz):
            d = direction(axes, f32(setup["az0"]) + f32(ia) * f32(setup["daz"]),
                          f32(setup["el0"]) + f32(ie) * f32(setup["del"]))
            for ir in range(nr):
                r = f32(f32(setup["r0"]) + f32(ir) * f32(setup["dr"]))
                P = (tx + r * d).astype(np.float32)
                if np.any(P < lo) or np.any(P > hi):
                    out[ir, ie, ia] = -1.0
                    continue
                outbound = tx_delay(P, tx, axes, tx_pos, tx_amp, float(f32(setup["wavelength"])))
                acc = 0j
                for i in range(n_el):
                    t = outbound + _len(P - np.asarray(element_pos[i], np.float32))
                    s = (t + 0.5 * dt) / sdt
                    if s < 0.0 or s >= m - 1:
                        continue
                    m0 = int(s)
                    f = f32(s - m0)
                    v = z[m0, i] * (f32(1.0) - f) + z[m0 + 1, i] * f
                    acc += float(weights[i]) * complex(v) * complex(math.cos(w * t), math.sin(w * t))
                out[ir, ie, ia] = abs(acc) ** 2
    return out
