"""Coaxial-line monitor (EM_Lab's CoaxMonitor, src/Coax.hpp, src/Coax.cpp) on the C ABI.

At every station k = 1 .. length-1 along the cable three quantities are measured after each step
(src/Coax.cpp:130-174): V, the E line integral from the inner conductor to the shield averaged
over the four sides and the inner conductor's node rows; I, the B loop integral around the inner
conductor; Icm, the loop integral around the whole shield (or the balun sleeve). The device keeps
them as running-DFT phasors at the drive frequency (`LineMonitor`). `CoaxMonitor.readings()`
fits the two waves V(t) = A e^{-j beta t} + B e^{+j beta t} to the stations by least squares,
with beta the lattice's own TEM propagation constant, and derives the far-end reflection, VSWR,
terminating impedance, forward power and common-mode level (src/Coax.cpp:216-288).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from .emfdtd import LineMonitor, SimulationGrid, coax_geometry_factor

ETA0 = 376.730313668


def _node_lo(c: float, size: int) -> int:
    # std::lround of a float (src/Coax.cpp:14): round half away from zero
    v = float(np.float32(c) - np.float32(0.5) * np.float32(size))
    return int(math.floor(abs(v) + 0.5)) * (1 if v >= 0 else -1)


def _node_hi(c: float, size: int) -> int:
    v = float(np.float32(c) + np.float32(0.5) * np.float32(size))
    return int(math.floor(abs(v) + 0.5)) * (1 if v >= 0 else -1)


@dataclass
class CoaxSetup:
    """CoaxSetup (src/Coax.hpp:31-47)."""
    name: str = "coax"
    axis: int = 0
    sign: int = 1
    start: int = 0
    center: Sequence[float] = (0.0, 0.0, 0.0)
    length: int = 0
    inner: int = 1
    outer: int = 5
    wall: int = 1
    sleeve_start: int = -1
    sleeve_end: int = -1
    sleeve_outer: int = 0
    epsilon_r: float = 1.0
    z0_ohm: float = 50.0
    wavelength: float = 16.0
    cell_size: float = 1e-3
    termination: str = "terminals"
    source_ohm: float = 50.0
    load_ohm: float = 50.0
    drive: float = 1.0
    pulsed: bool = False


@dataclass
class CoaxReading:
    """CoaxReading (src/Coax.hpp:49-65)."""
    valid: bool = False
    tV: List[float] = field(default_factory=list)
    tI: List[float] = field(default_factory=list)
    V: np.ndarray = field(default_factory=lambda: np.zeros(0, complex))
    I: np.ndarray = field(default_factory=lambda: np.zeros(0, complex))
    Icm: np.ndarray = field(default_factory=lambda: np.zeros(0, complex))
    v_now: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    i_now: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    cm_now: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    A: complex = 0j
    B: complex = 0j
    beta: float = 0.0
    z0_measured: float = 0.0
    gamma_end: complex = 0j
    z_end: complex = 0j
    vswr: float = 1.0
    return_loss_db: float = 0.0
    forward_fraction: float = 0.0
    fit_residual: float = 1.0
    common_mode_db: float = -300.0
    peak_now: float = 0.0


def build_terms(lines: Sequence[CoaxSetup]):
    """The quantity / term / weight lists of CoaxMonitor::configure (src/Coax.cpp:95-160), and per
    line (first quantity, stations)."""
    quantities, terms, weights, layout = [], [], [], []
    for c in lines:
        a = c.axis
        u, v = (a + 1) % 3, (a + 2) % 3

        def texel(ia, iu, iv):
            t = [0, 0, 0, 0]
            t[a], t[u], t[v] = ia, iu, iv
            return t

        def add(t, comp, w):
            t = list(t)
            t[3] = comp
            terms.append(t)
            weights.append(np.float32(w))

        def loop(ia, u0, u1, v0, v1, w):
            # current along +a through the node rectangle [u0, u1] x [v0, v1] (src/Coax.cpp:119-129)
            for iv in range(v0, v1 + 1):
                add(texel(ia, u1, iv), 3 + v, w)
                add(texel(ia, u0 - 1, iv), 3 + v, -w)
            for iu in range(u0, u1 + 1):
                add(texel(ia, iu, v1), 3 + u, -w)
                add(texel(ia, iu, v0 - 1), 3 + u, w)

        cu, cv = float(c.center[u]), float(c.center[v])
        iu0, iu1 = _node_lo(cu, c.inner), _node_hi(cu, c.inner)
        iv0, iv1 = _node_lo(cv, c.inner), _node_hi(cv, c.inner)
        su0, su1 = _node_lo(cu, c.outer), _node_hi(cu, c.outer)
        sv0, sv1 = _node_lo(cv, c.outer), _node_hi(cv, c.outer)
        shield = c.outer + 2 * c.wall
        first_q = len(quantities)
        for k in range(1, c.length):
            ia = c.start + c.sign * k
            first = len(terms)
            w = np.float32(1.0) / (np.float32(4.0) * np.float32(c.inner + 1))
            for r in range(iv0, iv1 + 1):
                for i in range(iu1, su1):
                    add(texel(ia, i, r), u, w)
                for i in range(su0, iu0):
                    add(texel(ia, i, r), u, -w)
            for r in range(iu0, iu1 + 1):
                for i in range(iv1, sv1):
                    add(texel(ia, r, i), v, w)
                for i in range(sv0, iv0):
                    add(texel(ia, r, i), v, -w)
            quantities.append([first, len(terms) - first, 0, 0])
            ib = c.start + k if c.sign > 0 else c.start - k - 1
            s = float(c.sign)
            f2 = len(terms)
            loop(ib, iu0, iu1, iv0, iv1, s)
            quantities.append([f2, len(terms) - f2, 1, 0])
            in_sleeve = c.sleeve_outer > 0 and c.sleeve_start <= k < c.sleeve_end
            out = c.sleeve_outer if in_sleeve else shield
            f2 = len(terms)
            loop(ib, _node_lo(cu, out), _node_hi(cu, out), _node_lo(cv, out), _node_hi(cv, out), s)
            quantities.append([f2, len(terms) - f2, 1, 0])
        layout.append((first_q, max(c.length - 1, 0)))
    q = np.array(quantities, np.int32).reshape(-1, 4)
    t = np.array(terms, np.int32).reshape(-1, 4)
    return q, t, np.array(weights, np.float32), layout


def lattice_beta(wavelength: float, epsilon_r: float, dt: float) -> float:
    """The lattice's TEM propagation constant (rad per cell): sin(beta / 2) = sqrt(eps_r)
    sin(omega dt / 2) / dt (src/Coax.cpp:242-245)."""
    w = 2.0 * math.pi / max(float(np.float32(wavelength)), 2.0)
    s_arg = math.sqrt(float(np.float32(epsilon_r))) * math.sin(0.5 * w * dt) / dt
    return 2.0 * math.asin(min(max(s_arg, -1.0), 1.0))


def analyse(c: CoaxSetup, stations: int, phasors: np.ndarray, now: np.ndarray, dt: float,
            peak: float = 0.0) -> CoaxReading:
    """CoaxMonitor::analyse (src/Coax.cpp:216-288) on one line's phasors (3 quantities per
    station, stage-2 components) and instantaneous values, in FP64."""
    r = CoaxReading()
    n = stations
    ph = np.asarray(phasors, np.float32).reshape(-1, 4).astype(np.float64)
    nw = np.asarray(now, np.float32)
    r.tV = [float(k + 1) for k in range(n)]
    r.tI = [float(k + 1) + 0.5 for k in range(n)]
    r.V = ph[0:3 * n:3, 2] + 1j * ph[0:3 * n:3, 3]
    r.I = ph[1:3 * n:3, 2] + 1j * ph[1:3 * n:3, 3]
    r.Icm = ph[2:3 * n:3, 2] + 1j * ph[2:3 * n:3, 3]
    r.v_now, r.i_now, r.cm_now = nw[0:3 * n:3].copy(), nw[1:3 * n:3].copy(), nw[2:3 * n:3].copy()
    vmax = float(np.max(np.abs(r.v_now.astype(np.float64)))) if n else 0.0
    r.peak_now = max(vmax, peak * 0.995)
    r.beta = lattice_beta(c.wavelength, c.epsilon_r, dt)
    if c.pulsed or n < 8:
        return r
    edge = float(c.outer + 2)

    def fit(t, x):
        m11 = m12 = m22 = r1 = r2 = 0j
        for tk, xk in zip(t, x):
            if tk < edge or tk > c.length - edge:
                continue
            e = complex(math.cos(-r.beta * tk), math.sin(-r.beta * tk))
            f = e.conjugate()
            m11 += abs(e) ** 2
            m12 += e.conjugate() * f
            m22 += abs(f) ** 2
            r1 += e.conjugate() * xk
            r2 += f.conjugate() * xk
        det = m11 * m22 - m12 * m12.conjugate()
        if abs(det) < 1e-30:
            return None
        return (r1 * m22 - m12 * r2) / det, (m11 * r2 - m12.conjugate() * r1) / det

    fv = fit(r.tV, r.V)
    fi = fit(r.tI, r.I) if fv is not None else None
    if fv is None or fi is None or abs(fv[0]) < 1e-20:
        return r
    r.A, r.B = fv
    ai = fi[0]
    res = tot = 0.0
    for tk, vk in zip(r.tV, r.V):
        if tk < edge or tk > c.length - edge:
            continue
        model = r.A * complex(math.cos(-r.beta * tk), math.sin(-r.beta * tk)) + \
            r.B * complex(math.cos(r.beta * tk), math.sin(r.beta * tk))
        res += abs(vk - model) ** 2
        tot += abs(model) ** 2
    r.fit_residual = math.sqrt(res / tot) if tot > 0.0 else 1.0
    r.z0_measured = abs(r.A / ai) * ETA0 if abs(ai) > 1e-30 else 0.0
    ang = 2.0 * r.beta * c.length
    r.gamma_end = r.B / r.A * complex(math.cos(ang), math.sin(ang))
    g = min(abs(r.gamma_end), 0.999999)
    r.vswr = (1.0 + g) / (1.0 - g)
    r.return_loss_db = -20.0 * math.log10(g) if g > 1e-12 else 200.0
    z0 = float(np.float32(c.z0_ohm))
    r.z_end = z0 * (1.0 + r.gamma_end) / (1.0 - r.gamma_end)
    z0n, rsn = z0 / ETA0, float(np.float32(c.source_ohm)) / ETA0
    drive = float(np.float32(c.drive))
    p_fwd, p_avail = abs(r.A) ** 2 / (2.0 * z0n), drive * drive / (8.0 * rsn)
    r.forward_fraction = p_fwd / p_avail if p_avail > 0.0 else 0.0
    i_fwd = abs(r.A) / z0n
    cm = float(np.max(np.abs(r.Icm))) if n else 0.0
    r.common_mode_db = 20.0 * math.log10(cm / i_fwd) if cm > 0.0 and i_fwd > 0.0 else -300.0
    r.valid = True
    return r


class CoaxMonitor:
    """CoaxMonitor (src/Coax.hpp:67-104) over the grid's line monitor."""

    def __init__(self, grid: SimulationGrid):
        self.grid = grid
        self.lines: List[CoaxSetup] = []
        self.layout: list = []
        self.averaging = 4.0
        self.start_time = 0.0
        self._dev = LineMonitor(grid)
        self._peaks: List[float] = []
        self._pending = None
        self.polled: List[CoaxReading] = []

    def configure(self, lines: Sequence[CoaxSetup], averaging_periods: float = 4.0) -> None:
        self.lines = list(lines)
        self.averaging = max(float(averaging_periods), 0.5)
        self.start_time = self.grid.time
        q, t, w, self.layout = build_terms(self.lines)
        wl = self.lines[0].wavelength if self.lines else 16.0
        self._dev.configure(q, t, w, wl, self.averaging)
        self._peaks = [0.0] * len(self.lines)
        self._pending = None
        self.polled = [CoaxReading() for _ in self.lines]

    def restart(self) -> None:
        self.configure(self.lines, self.averaging)

    def disable(self) -> None:
        self.configure([], self.averaging)

    def enabled(self) -> bool:
        return bool(self.lines)

    def periods(self) -> float:
        if not self.lines:
            return 0.0
        return max(self.grid.time - self.start_time, 0.0) / max(self.lines[0].wavelength, 2.0)

    def poll(self) -> None:
        """CoaxMonitor::poll (src/Coax.cpp:199-214), once per frame: the readback taken at the
        previous poll is analysed into `polled`, then the current phasors are read back."""
        if self._pending is not None:
            self.polled = self._analyse(*self._pending)
        self._pending = self._dev.read()

    def readings(self) -> List[CoaxReading]:
        """The analysis of the phasors as they are now."""
        return self._analyse(*self._dev.read())

    def _analyse(self, ph, now) -> List[CoaxReading]:
        out = []
        for i, c in enumerate(self.lines):
            first, stations = self.layout[i]
            r = analyse(c, stations, ph[first:first + 3 * stations], now[first:first + 3 * stations], self.grid.dt,
                        self._peaks[i])
            self._peaks[i] = r.peak_now
            out.append(r)
        return out

    @staticmethod
    def geometry_factor(inner: int, outer: int) -> float:
        return coax_geometry_factor(inner, outer)
