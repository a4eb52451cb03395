"""Scene description -> solver inputs (materials, sub-cell model, sources, ports, monitors).

A Python counterpart of the numerically relevant part of EM_Lab's Lua scene API
(src/ScriptEngine.cpp) and EnvironmentBuilder (src/Environment.cpp): material resolution,
voxel fill, thin-wire dipoles with resistive feeds, ideal (Hertzian) dipoles, thin wires with
the free-end correction, the far-field box and the sweep band. The arithmetic keeps the
reference's float/double boundaries so the sub-cell coefficients match bit for bit.

Also holds the seeded generators of the benchmark scenes (SplitMix64 -> xoshiro256**).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .emfdtd import (WAVE_CW, WAVE_PULSE, FarFieldSetup, FluxSurface, GridConfig, PortSetup, SourceDesc,
                     SubcellModel, SweepSetup)

f32 = np.float32
ETA0 = 376.730313668            # src/ScriptEngine.cpp:27
ETA0F = f32(376.730313668)      # src/ScriptEngine.cpp:28
C0 = 299792458.0
PERFECT_EDGE_SIGMA = f32(6.0e4)  # src/SimulationGrid.hpp:103
MAX_COURANT = float(f32(0.99) / f32(1.7320508))  # src/SimulationGrid.cpp:24


def round_half_away(x: float) -> int:
    """glm::round / std::lround: halves away from zero."""
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


# ---------------------------------------------------------------------------------------------
# PRNG: SplitMix64 seeding xoshiro256** (Blackman & Vigna)
# ---------------------------------------------------------------------------------------------
_M64 = (1 << 64) - 1


class Xoshiro256ss:
    def __init__(self, seed: int):
        s = seed & _M64
        self.s = []
        for _ in range(4):
            s = (s + 0x9E3779B97F4A7C15) & _M64
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
            self.s.append(z ^ (z >> 31))

    @staticmethod
    def _rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & _M64

    def next(self) -> int:
        s = self.s
        result = (self._rotl((s[1] * 5) & _M64, 7) * 9) & _M64
        t = (s[1] << 17) & _M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return result

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def log_uniform(self, lo: float, hi: float) -> float:
        return math.exp(self.uniform(math.log(lo), math.log(hi)))


# ---------------------------------------------------------------------------------------------
# Scene
# ---------------------------------------------------------------------------------------------
@dataclass
class ThinWireSpec:
    node: tuple
    axis: int
    i0: int
    i1: int
    radius: float
    gap_edge: Optional[int] = None


@dataclass
class Scene:
    """Settings as scene{} (src/ScriptEngine.cpp:605-650) and the EnvironmentBuilder voxels."""
    size: tuple = (128, 128, 128)
    pml: int = 16
    courant: float = 0.5
    cell_size: float = 1.0e-3
    wavelength: float = 16.0
    pml_sigma_max: float = 1.6
    pml_grading: float = 3.0
    sigma: np.ndarray = None
    eps: np.ndarray = None
    subcell: SubcellModel = field(default_factory=SubcellModel)
    sources: list = field(default_factory=list)
    ports: list = field(default_factory=list)
    thin_wires: list = field(default_factory=list)
    sweep: Optional[SweepSetup] = None
    far_field: Optional[FarFieldSetup] = None
    fluxes: list = field(default_factory=list)
    warnings: list = field(default_factory=list)
    coaxes: list = field(default_factory=list)
    # radar targets (`target = true` objects, src/ScriptEngine.cpp:830-838): their TargetInfo, and
    # the materials without them (the radar's calibration scene), kept once a target is placed
    targets: list = field(default_factory=list)
    background: Optional[tuple] = None
    emitters: list = field(default_factory=list)  # emitter{} transmitters (EmitterSetup)
    radar: Optional[object] = None  # radar{} (RadarSetup)
    crpa: Optional[object] = None   # crpa{} (CrpaSetup)

    def __post_init__(self):
        nx, ny, nz = self.size
        if self.sigma is None:
            self.sigma = np.zeros((nz, ny, nx), np.float32)
        if self.eps is None:
            self.eps = np.ones((nz, ny, nx), np.float32)

    # -- grid --
    def grid_config(self) -> GridConfig:
        return GridConfig(tuple(self.size), float(self.courant), int(self.pml), float(self.pml_sigma_max),
                          float(self.pml_grading))

    # -- materials (src/ScriptEngine.cpp:458-465) --
    def resolve(self, eps_r: float = 1.0, sigma: Optional[float] = None, conductivity: Optional[float] = None,
                loss_tangent: Optional[float] = None):
        s = 0.0
        if sigma is not None:
            s = sigma
        elif conductivity is not None:
            s = conductivity * float(f32(self.cell_size)) * ETA0
        elif loss_tangent is not None:
            s = loss_tangent * (2.0 * math.pi / float(f32(self.wavelength))) * float(f32(eps_r))
        return f32(s), max(f32(eps_r), f32(1.0))

    def _set(self, x, y, z, sig, eps):
        # MaterialVolume::set (src/SimulationGrid.cpp:42-47)
        self.sigma[z, y, x] = max(sig, f32(0.0))
        self.eps[z, y, x] = max(eps, f32(1.0))

    def _paint(self, idx, sig, eps, target: bool, name: str) -> None:
        """Fills cells (index arrays z, y, x); a radar target goes into the full scene only and
        is recorded (targetInfo, src/ScriptEngine.cpp:800-838), other objects into both."""
        if target:
            from .radar import target_info
            if self.background is None:
                self.background = (self.sigma.copy(), self.eps.copy())
            mask = np.zeros(self.sigma.shape, bool)
            mask[idx] = True
            self.targets.append(target_info(name, mask, self.size))
        elif self.background is not None:
            self.background[0][idx] = max(sig, f32(0.0))
            self.background[1][idx] = max(eps, f32(1.0))
        self.sigma[idx] = max(sig, f32(0.0))
        self.eps[idx] = max(eps, f32(1.0))

    def background_materials(self):
        """(sigma, eps) flattened without the radar targets (the calibration scene)."""
        if self.background is None:
            return self.flat_materials()
        return np.ascontiguousarray(self.background[0].ravel()), np.ascontiguousarray(self.background[1].ravel())

    def sphere(self, center, radius, eps_r=1.0, sigma=None, conductivity=None, loss_tangent=None, target=False,
               name="sphere") -> int:
        """EnvironmentBuilder::sphere -> fill(EllipsoidShape) (src/Environment.cpp:82-105)."""
        sig, eps = self.resolve(eps_r, sigma, conductivity, loss_tangent)
        c = np.asarray(center, np.float32)
        r = max(f32(radius), f32(1e-3))
        nx, ny, nz = self.size
        lo = [min(max(int(math.floor(float(c[a] - r - f32(0.5)))), 0), self.size[a]) for a in range(3)]
        hi = [min(max(int(math.ceil(float(c[a] + r + f32(0.5)))), 0), self.size[a]) for a in range(3)]
        zz, yy, xx = np.meshgrid(np.arange(lo[2], hi[2]), np.arange(lo[1], hi[1]), np.arange(lo[0], hi[0]),
                                 indexing="ij")
        px = (xx.astype(np.float32) + f32(0.5) - c[0]) / r
        py = (yy.astype(np.float32) + f32(0.5) - c[1]) / r
        pz = (zz.astype(np.float32) + f32(0.5) - c[2]) / r
        inside = (px * px + py * py) + pz * pz <= f32(1.0)
        self._paint((zz[inside], yy[inside], xx[inside]), sig, eps, target, name)
        return int(inside.sum())

    def box(self, mn, mx, eps_r=1.0, sigma=None, conductivity=None, target=False, name="box") -> int:
        """EnvironmentBuilder::box (src/Environment.cpp:99-101): cells with min <= centre < max."""
        sig, eps = self.resolve(eps_r, sigma, conductivity)
        mn = np.minimum(np.asarray(mn, np.float32), np.asarray(mx, np.float32))
        mx2 = np.maximum(np.asarray(mn, np.float32), np.asarray(mx, np.float32))
        idx = []
        for a in range(3):
            lo = min(max(int(math.floor(float(mn[a] - f32(0.5)))), 0), self.size[a])
            hi = min(max(int(math.ceil(float(mx2[a] + f32(0.5)))), 0), self.size[a])
            cc = np.arange(lo, hi).astype(np.float32) + f32(0.5)
            idx.append(np.arange(lo, hi)[(cc >= mn[a]) & (cc < mx2[a])])
        if any(len(i) == 0 for i in idx):
            return 0
        self._paint(np.ix_(idx[2], idx[1], idx[0]), sig, eps, target, name)
        return len(idx[0]) * len(idx[1]) * len(idx[2])

    # -- sources --
    def waveform(self, wavelength=None, waveform="cw", ramp=3.0, pulse_width=1.0, phase_deg=0.0, interval=0.0,
                 delay=0.0):
        """readWaveform (src/ScriptEngine.cpp:1082-1102); the scene wavelength is the default."""
        if wavelength is None:
            wavelength = self.wavelength
        s = SourceDesc(wavelength=float(f32(wavelength)), phase=float(f32(math.radians(phase_deg))))
        if waveform in ("cw", "continuous", "sine"):
            s.waveform = WAVE_CW
            s.ramp_periods = float(f32(ramp))
        else:
            s.waveform = WAVE_PULSE
            s.ramp_periods = float(f32(pulse_width))
        s.repeat_periods = float(f32(interval))
        s.delay_periods = float(f32(delay))
        return s

    def hertzian(self, center, orientation=(0, 0, 1), amplitude=1.0, length=0.0, **wave):
        """dipole{model="current"}: SourceDesc::dipole + copyTiming (src/ScriptEngine.cpp:1314-1321)."""
        w = self.waveform(**wave)
        soft = f32(0.75)
        amp = f32(amplitude) / (f32(math.pi) * soft * soft)
        ax = np.asarray(orientation, np.float32)
        len2 = f32(np.dot(ax, ax))
        n = ax / np.sqrt(len2) if len2 > 0 else np.array([0, 0, 1], np.float32)
        half = np.abs(n) * (f32(0.5) * f32(max(length, 0.0)))
        s = SourceDesc(center=tuple(float(v) for v in np.asarray(center, np.float32)),
                       half_extent=tuple(float(v) for v in half), softness=(0.75, 0.75, 0.75),
                       polarization=tuple(float(v) for v in ax), amplitude=float(amp))
        for k in ("wavelength", "phase", "waveform", "ramp_periods", "repeat_periods", "delay_periods", "enabled"):
            setattr(s, k, getattr(w, k))
        self.sources.append(s)
        return s

    def add_thin_wire(self, node, axis, i0, i1, radius, gap_edge=None):
        """addThinWire (src/ScriptEngine.cpp:1183-1200)."""
        if i1 <= i0:
            return
        self.thin_wires.append(ThinWireSpec(tuple(int(v) for v in node), axis, i0, i1, float(f32(radius)), gap_edge))

    def add_feed(self, wave: SourceDesc, texel, axis, sign, edges, impedance=50.0, drive=1.0, voltage=True,
                 name="dipole"):
        """addFeed (src/ScriptEngine.cpp:1205-1252)."""
        u, v = (axis + 1) % 3, (axis + 2) % 3
        s = SourceDesc(**{k: getattr(wave, k) for k in wave.__dataclass_fields__})
        s.softness = (0.25, 0.25, 0.25)
        pol = [0.0, 0.0, 0.0]
        pol[axis] = float(sign)
        s.polarization = tuple(pol)
        if edges == 4:
            s.center = tuple(float(f32(t) + f32(0.5)) for t in texel)
            he = [0.5, 0.5, 0.5]
            he[axis] = 0.0
            s.half_extent = tuple(he)
        else:
            c = [float(t) for t in texel]
            c[axis] = float(f32(c[axis]) + f32(0.5))
            s.center = tuple(c)
            s.half_extent = (0.0, 0.0, 0.0)
        if voltage:
            sigma = ETA0F / (f32(edges) * f32(impedance))
            offs = [(0, 0, 0)]
            eu = [0, 0, 0]
            ev = [0, 0, 0]
            eu[u] = 1
            ev[v] = 1
            offs = [(0, 0, 0), tuple(eu), tuple(ev), tuple(a + b for a, b in zip(eu, ev))]
            for i in range(edges):
                self.subcell.edges.append((tuple(t + o for t, o in zip(texel, offs[i])), axis, float(sigma)))
            s.amplitude = float(-sigma * f32(drive))
        else:
            s.amplitude = float(f32(drive) / f32(edges))
        s.port = len(self.ports)
        if s.enabled:
            self.sources.append(s)
        self.ports.append(PortSetup(name=name, texel=tuple(int(t) for t in texel), axis=axis, sign=float(sign),
                                    edges=edges, wavelength=wave.wavelength, voltage_feed=voltage,
                                    impedance_ohm=float(f32(impedance)), drive=float(f32(drive))))

    def coax(self, start, axis=0, sign=1, length=None, outer=5, wall=1, z0=50.0, eps_r=None, termination="load",
             load=None, balun=False, balun_length=None, balun_gap=2, amplitude=1.0, impedance=None, name="cable",
             **wave):
        """api_coax (src/ScriptEngine.cpp:1961-2127): a square coaxial cable along `axis` (0-2,
        `sign` toward the far end) from the lattice plane through `start` (a point; its
        transverse coordinates pick the cell-centre line). Shield (with end plates unless the
        termination is "terminals"), a dielectric giving `z0` from the lattice geometry factor
        (or `eps_r` as given), a 1x1-cell inner conductor, a generator across the gap cell at the
        start, a resistor across the last cell for termination "load" (also "open", "short",
        "terminals"), an optional quarter-wave sleeve balun; registers the line monitor's
        CoaxSetup in `coaxes`. Returns (CoaxSetup, inner terminal node, outer terminal node)."""
        from .coax import CoaxSetup
        from .emfdtd import coax_geometry_factor
        wv = self.waveform(**wave)
        ax, s = int(axis), (1 if sign >= 0 else -1)
        u, v = (ax + 1) % 3, (ax + 2) % 3
        ln = round_half_away(4.0 * wv.wavelength if length is None else float(length))
        inner = 1
        if outer < 3 or outer % 2 == 0:
            raise ValueError("coax: outer (the inside of the shield, cells) must be odd and >= 3")
        if wall < 1:
            raise ValueError("coax: wall must be >= 1 cell")
        term = termination.lower()
        if term not in ("load", "open", "short", "terminals"):
            raise ValueError('coax: termination must be "load", "open", "short" or "terminals"')
        if ln < 2 * outer + 12:
            raise ValueError(f"coax: length must be at least {2 * outer + 12} cells for this cross-section")
        g = coax_geometry_factor(inner, outer)
        z0_air = ETA0 / g
        if eps_r is None:
            er = (z0_air / z0) * (z0_air / z0)
            if er < 1.0:
                self.warnings.append(f"coax: {z0:.1f} ohm needs a wider shield; with outer = {outer} the "
                                     f"air-filled line is {z0_air:.1f} ohm")
                er = 1.0
            diel_eps = f32(er)
        else:
            diel_eps = max(f32(eps_r), f32(1.0))
        z0_line = z0_air / math.sqrt(max(float(diel_eps), 1.0))
        st = round_half_away(float(start[ax]))
        centre = [0.0, 0.0, 0.0]
        centre[ax] = float(st)
        centre[u] = math.floor(float(f32(start[u]))) + 0.5
        centre[v] = math.floor(float(f32(start[v]))) + 0.5

        def put(t0, t1, size, eps=1.0, sigma=0.0):
            lo, hi = [0.0] * 3, [0.0] * 3
            lo[ax] = float(st + t0 if s > 0 else st - t1)
            hi[ax] = float(st + t1 if s > 0 else st - t0)
            for k in (u, v):
                lo[k] = float(f32(centre[k]) - f32(0.5) * f32(size))
                hi[k] = float(f32(centre[k]) + f32(0.5) * f32(size))
            self.box(lo, hi, eps_r=eps, sigma=sigma)

        metal = float(f32(4.0) * PERFECT_EDGE_SIGMA)  # every edge touching it reaches the clamp
        shield = outer + 2 * wall
        capped = term != "terminals"
        c = CoaxSetup(name=name, axis=ax, sign=s, start=st, center=tuple(centre), length=ln, inner=inner, outer=outer,
                      wall=wall, epsilon_r=float(diel_eps), z0_ohm=float(f32(z0_line)), wavelength=wv.wavelength,
                      cell_size=float(f32(self.cell_size)), termination=term, pulsed=wv.waveform == WAVE_PULSE)
        if balun:
            bl = round_half_away(0.25 * wv.wavelength if balun_length is None else float(balun_length))
            gap = round_half_away(float(balun_gap))
            sleeve = shield + 2 * (gap + wall)
            if bl + wall >= ln - 4:
                raise ValueError("coax: balun_length is too long for this cable")
            put(ln - bl - wall, ln, sleeve, sigma=metal)
            put(ln - bl, ln, shield + 2 * gap)
            c.sleeve_start, c.sleeve_end, c.sleeve_outer = ln - bl - wall, ln, sleeve
        put(-wall, ln + wall if capped else ln, shield, sigma=metal)
        put(0, ln, outer, eps=float(diel_eps))
        inner_end = {"terminals": ln + 3, "short": ln, "load": ln - 1}.get(term, ln - 2)
        put(1, inner_end, inner, sigma=metal)
        gen_ohm = float(f32(z0_line if impedance is None else impedance))
        gap_cell = [0, 0, 0]
        gap_cell[ax] = st if s > 0 else st - 1
        gap_cell[u] = int(math.floor(centre[u]))
        gap_cell[v] = int(math.floor(centre[v]))
        self.add_feed(wv, tuple(gap_cell), ax, float(s), 4, impedance=gen_ohm, drive=amplitude, voltage=True,
                      name=name)
        c.source_ohm, c.drive = gen_ohm, float(f32(amplitude))
        c.load_ohm = float(f32(z0_line if load is None else load))
        if term == "load":
            quiet = SourceDesc(**{k: getattr(wv, k) for k in wv.__dataclass_fields__})
            quiet.enabled = True
            cell = list(gap_cell)
            cell[ax] = st + ln - 1 if s > 0 else st - ln
            self.add_feed(quiet, tuple(cell), ax, float(-s), 4, impedance=c.load_ohm, drive=0.0, voltage=True,
                          name=name + " load")
        elif load is not None:
            self.warnings.append("coax: 'load' only applies to termination = \"load\"")
        inner_node = outer_node = None
        if term == "terminals":
            tax = 1 if ax == 2 else 2
            ts = 1
            tu = v if tax == u else u
            lo, hi = [0.0] * 3, [0.0] * 3
            lo[ax] = float(st + ln if s > 0 else st - ln - 3)
            hi[ax] = float(st + ln + 3 if s > 0 else st - ln)
            lo[tu], hi[tu] = centre[tu] - 0.5, centre[tu] + 0.5
            wall_in = float(f32(0.5) * f32(outer))
            wall_out = wall_in + float(wall)
            lo[tax] = centre[tax] - wall_out if ts > 0 else centre[tax] + wall_in
            hi[tax] = centre[tax] - wall_in if ts > 0 else centre[tax] + wall_out
            self.box(lo, hi, sigma=metal)
            inner_node, outer_node = [0.0] * 3, [0.0] * 3
            inner_node[ax] = outer_node[ax] = float(st + s * (ln + 2))
            inner_node[tu] = outer_node[tu] = centre[tu] - 0.5
            inner_node[tax] = centre[tax] + 0.5 * ts
            outer_node[tax] = centre[tax] - ts * wall_out
        self.coaxes.append(c)
        return c, inner_node, outer_node

    def dipole_thin(self, center, axis=2, length=11.0, radius=None, impedance=50.0, amplitude=1.0, sign=1.0,
                    name="dipole", **wave):
        """dipole{model="thin"} (src/ScriptEngine.cpp:1255-1313): arms, thin wire, resistive feed."""
        w = self.waveform(**wave)
        if length < 5.0:
            raise ValueError("a thin-wire dipole needs length >= 5 cells")
        arm = max(1, round_half_away((float(f32(length)) - 1.0) * 0.5))
        if abs(2 * arm + 1 - length) >= 1.0:
            self.warnings.append(f"dipole: length {length:.1f} realised as {2 * arm + 1} cells")
        if radius is None:
            radius = float(f32(w.wavelength) / f32(400.0))
        c = np.asarray(center, np.float32)
        node = [round_half_away(float(v)) for v in c]
        node[axis] = int(math.floor(float(c[axis])))
        g = node[axis]
        self.add_thin_wire(node, axis, g - arm, g + 1 + arm, radius, g)
        self.add_feed(w, tuple(node), axis, sign, 1, impedance, amplitude, True, name)

    def emitter(self, position, name=None, polarization="linear", toward=None, amplitude=1.0, **wave):
        """emitter{} (src/ScriptEngine.cpp:1850-1892): a point transmitter of ideal dipoles,
        linear (a direction) or circular ("rhcp" / "lhcp", toward the grid centre by default)."""
        from .crpa import EmitterSetup, emitter_sources
        w = self.waveform(**wave)
        name = name or f"emitter {len(self.emitters) + 1}"
        if any(e.name == name for e in self.emitters):
            raise ValueError(f"an emitter named '{name}' already exists")
        center = tuple(float(f32(v) * f32(0.5)) for v in self.size)
        srcs = emitter_sources(len(self.emitters), position, w, polarization, toward, amplitude, center)
        self.sources.extend(srcs)
        pol = polarization.lower() if isinstance(polarization, str) else "linear"
        self.emitters.append(EmitterSetup(name, tuple(float(v) for v in position), pol))

    def wire(self, p0, p1, radius=None):
        """wire{} (src/ScriptEngine.cpp:1380-1394)."""
        a = [round_half_away(float(v)) for v in np.asarray(p0, np.float32)]
        b = [round_half_away(float(v)) for v in np.asarray(p1, np.float32)]
        d = [bb - aa for aa, bb in zip(a, b)]
        if sum(1 for v in d if v != 0) != 1:
            raise ValueError("from and to must differ along exactly one axis")
        ax = next(i for i in range(3) if d[i] != 0)
        if radius is None:
            radius = float(f32(self.wavelength) / f32(400.0))
        self.add_thin_wire(a, ax, min(a[ax], b[ax]), max(a[ax], b[ax]), radius)

    # -- monitors --
    def set_far_field(self, margin=4, wavelength=None, averaging=4.0, axis=(0, 0, 1)):
        """far_field{} default box (src/ScriptEngine.cpp:2160-2189)."""
        lo = tuple(self.pml + 1 + margin for _ in range(3))
        hi = tuple(s - self.pml - 2 - margin for s in self.size)
        self.far_field = FarFieldSetup(lo, hi, float(f32(wavelength or self.wavelength)), float(f32(averaging)),
                                       tuple(float(v) for v in axis))
        return self.far_field

    def set_sweep(self, wavelengths, points=81):
        """sweep{wavelengths=...} (src/ScriptEngine.cpp:1431-1454)."""
        lo, hi = min(wavelengths), max(wavelengths)
        self.sweep = SweepSetup(float(f32(lo)), float(f32(hi)), int(min(max(points, 2), 2001)))
        return self.sweep

    # -- emission (src/ScriptEngine.cpp:2346-2419) --
    def emit_thin_wires(self) -> None:
        def key(p):
            return tuple(int(v) for v in p)

        nodes = []
        for t in self.thin_wires:
            s = set()
            for k in range(t.i0, t.i1 + 1):
                p = list(t.node)
                p[t.axis] = k
                s.add(key(p))
            nodes.append(s)
        sheet_nodes = set()
        for (tex, ax, sg) in self.subcell.edges:
            if f32(sg) < f32(0.5) * PERFECT_EDGE_SIGMA:
                continue
            end = list(tex)
            end[ax] += 1
            sheet_nodes.add(key(tex))
            sheet_nodes.add(key(end))
        nx, ny, nz = self.size

        def mat_sigma(p):
            x, y, z = p
            if 0 <= x < nx and 0 <= y < ny and 0 <= z < nz:
                return self.sigma[z, y, x]
            return f32(0.0)

        def touches(self_i, p):
            if key(p) in sheet_nodes:
                return True
            for w, s in enumerate(nodes):
                if w != self_i and key(p) in s:
                    return True
            for dz in (-1, 0):
                for dy in (-1, 0):
                    for dx in (-1, 0):
                        if mat_sigma((p[0] + dx, p[1] + dy, p[2] + dz)) >= f32(1.0):
                            return True
            return False

        for w, t0 in enumerate(self.thin_wires):
            t = ThinWireSpec(**t0.__dict__)
            a = list(t.node)
            b = list(t.node)
            a[t.axis] = t.i0
            b[t.axis] = t.i1
            gap = t.gap_edge if t.gap_edge is not None else -(1 << 31)
            freeA = (not touches(w, a)) and gap != t.i0
            freeB = (not touches(w, b)) and gap != t.i1 - 1
            keep = 2 if t.i0 <= gap < t.i1 else 1
            if (t.i1 - t.i0) - int(freeA) - int(freeB) >= keep:
                if freeA:
                    t.i0 += 1
                if freeB:
                    t.i1 -= 1
            elif freeA or freeB:
                self.warnings.append(f"thin wire at {tuple(a)} is too short for the end correction")
            fac = float(f32(2.0) / np.log(f32(1.0) / f32(t.radius)))
            u, v = (t.axis + 1) % 3, (t.axis + 2) % 3
            eu = [0, 0, 0]
            ev = [0, 0, 0]
            eu[u] = 1
            ev[v] = 1

            def term(comp):
                return 0 if (comp + 2) % 3 == t.axis else 1

            for k in range(t.i0, t.i1):
                q = list(t.node)
                q[t.axis] = k
                if k != gap:
                    self.subcell.edges.append((tuple(q), t.axis, float(PERFECT_EDGE_SIGMA)))
                qmv = tuple(qq - e for qq, e in zip(q, ev))
                qmu = tuple(qq - e for qq, e in zip(q, eu))
                self.subcell.wire_terms.append((tuple(q), u, term(u), fac))
                self.subcell.wire_terms.append((qmv, u, term(u), fac))
                self.subcell.wire_terms.append((tuple(q), v, term(v), fac))
                self.subcell.wire_terms.append((qmu, v, term(v), fac))
        self.thin_wires = []

    def flat_materials(self):
        return np.ascontiguousarray(self.sigma.ravel()), np.ascontiguousarray(self.eps.ravel())


# ---------------------------------------------------------------------------------------------
# Benchmark / parity scenes
# ---------------------------------------------------------------------------------------------
DEFAULT_SEED = 20251203


def hertzian_vacuum(n=64, pml=8, steps_wavelength=16.0, courant=MAX_COURANT) -> Scene:
    """Plain FDTD: n^3 vacuum, PML, z Hertzian dipole at the centre, Gaussian pulse (lambda 16)."""
    sc = Scene(size=(n, n, n), pml=pml, courant=courant, wavelength=steps_wavelength)
    sc.hertzian((n / 2, n / 2, n / 2), (0, 0, 1), 1.0, wavelength=steps_wavelength, waveform="pulse", pulse_width=1.0)
    sc.set_far_field(wavelength=steps_wavelength)
    return sc


def spheres_dipole(n=256, pml=10, n_spheres=32, wavelength=20.0, seed=DEFAULT_SEED, length=11,
                   radius_range=(4.0, 20.0), sweep_points=32) -> Scene:
    """n^3 grid, seeded random dielectric spheres and a z thin-wire half-wave dipole fed by a
    50 ohm resistive voltage source (pulse); sweep, flux box and far-field monitors."""
    sc = Scene(size=(n, n, n), pml=pml, courant=MAX_COURANT, cell_size=1.0e-3, wavelength=wavelength)
    ff = sc.set_far_field(wavelength=wavelength)
    rng = Xoshiro256ss(seed)
    c0 = n / 2
    dip_lo, dip_hi = c0 - length / 2 - 1, c0 + length / 2 + 1
    placed = 0
    attempts = 0
    scale = n / 256.0
    while placed < n_spheres and attempts < 100000:
        attempts += 1
        r = rng.uniform(radius_range[0] * scale, radius_range[1] * scale)
        eps = rng.uniform(1.0, 12.0)
        cond = rng.log_uniform(1e-4, 1.0)
        lo = ff.box_min[0] + r + 2
        hi = ff.box_max[0] - r - 2
        if hi <= lo:
            continue
        c = (rng.uniform(lo, hi), rng.uniform(lo, hi), rng.uniform(lo, hi))
        # distance from the dipole's axis segment minus the radius must be >= 3 cells
        dz = max(dip_lo - c[2], 0.0, c[2] - dip_hi)
        dist = math.sqrt((c[0] - c0) ** 2 + (c[1] - c0) ** 2 + dz * dz)
        if dist - r < 3.0:
            continue
        sc.sphere(c, r, eps_r=eps, conductivity=cond)
        placed += 1
    radius = 0.005 * wavelength
    sc.dipole_thin((c0, c0, c0), axis=2, length=length, radius=radius, impedance=50.0, amplitude=1.0,
                   wavelength=wavelength, waveform="pulse", pulse_width=1.0)
    sc.set_sweep((wavelength / 1.5, wavelength / 0.5), points=sweep_points)
    m = ff.box_min[0] - 2
    sc.fluxes.append(FluxSurface.box((m, m, m), (n - 1 - m, n - 1 - m, n - 1 - m), "flux box"))
    sc.emit_thin_wires()
    return sc


def spheres_array(n=192, pml=10, n_spheres=32, wavelength=20.0, seed=DEFAULT_SEED, length=11, rows=4,
                  cols=4, sweep_points=32) -> Scene:
    """Precision-study scene: the seeded sphere medium of spheres_dipole (radii and placement
    scaled by n / 256) around a rows x cols array of z thin-wire dipoles at half-wave spacing in
    x and y, each on its own 50 ohm resistive feed (uniform amplitude, Gaussian pulse); port
    sweep and far-field box as spheres_dipole."""
    sc = Scene(size=(n, n, n), pml=pml, courant=MAX_COURANT, cell_size=1.0e-3, wavelength=wavelength)
    ff = sc.set_far_field(wavelength=wavelength)
    rng = Xoshiro256ss(seed)
    c0 = n / 2
    spacing = round_half_away(wavelength / 2)
    xs = [round_half_away(c0 + (i - (cols - 1) / 2) * spacing) for i in range(cols)]
    ys = [round_half_away(c0 + (j - (rows - 1) / 2) * spacing) for j in range(rows)]
    keep_lo = (min(xs) - 3, min(ys) - 3, c0 - length / 2 - 1 - 3)
    keep_hi = (max(xs) + 3, max(ys) + 3, c0 + length / 2 + 1 + 3)
    scale = n / 256.0
    placed = attempts = 0
    while placed < n_spheres and attempts < 100000:
        attempts += 1
        r = rng.uniform(4.0 * scale, 20.0 * scale)
        eps = rng.uniform(1.0, 12.0)
        cond = rng.log_uniform(1e-4, 1.0)
        lo = ff.box_min[0] + r + 2
        hi = ff.box_max[0] - r - 2
        if hi <= lo:
            continue
        c = (rng.uniform(lo, hi), rng.uniform(lo, hi), rng.uniform(lo, hi))
        # clear of the array's bounding box (plus 3 cells)
        d2 = sum(max(keep_lo[a] - c[a], 0.0, c[a] - keep_hi[a]) ** 2 for a in range(3))
        if math.sqrt(d2) - r < 0.0:
            continue
        sc.sphere(c, r, eps_r=eps, conductivity=cond)
        placed += 1
    for j, y in enumerate(ys):
        for i, x in enumerate(xs):
            sc.dipole_thin((x, y, c0), axis=2, length=length, radius=0.005 * wavelength, impedance=50.0,
                           amplitude=1.0, name=f"element {j * cols + i}", wavelength=wavelength, waveform="pulse",
                           pulse_width=1.0)
    sc.set_sweep((wavelength / 1.5, wavelength / 0.5), points=sweep_points)
    m = ff.box_min[0] - 2
    sc.fluxes.append(FluxSurface.box((m, m, m), (n - 1 - m, n - 1 - m, n - 1 - m), "flux box"))
    sc.emit_thin_wires()
    return sc


# ---------------------------------------------------------------------------------------------
# Batched antenna scenes: parallel z thin-wire elements in vacuum, element 0 fed
# ---------------------------------------------------------------------------------------------
@dataclass
class ArrayGeometry:
    """Realised geometry of a parallel-wire antenna, in cells: lengths[i] of element i,
    spacings[i] = x distance from element i-1 (spacings[0] = 0), radii[i]."""
    lengths: list
    spacings: list
    radii: list

    @property
    def elements(self) -> int:
        return len(self.lengths)

    def span(self) -> int:
        return int(sum(self.spacings))


ARRAY_LAMBDA0 = 16.0
ARRAY_PML = 8
ARRAY_MAX_ELEMENTS = 6


def draw_array_geometry(rng: "Xoshiro256ss", lambda0: float = ARRAY_LAMBDA0) -> ArrayGeometry:
    """1-6 parallel z wires: length U(0.3, 0.55) lambda0 (element 0 realised as an odd dipole of
    at least 5 cells, the others as whole cells), spacing U(0.1, 0.4) lambda0 (whole cells, at
    least 1), radius log-U(0.001, 0.01) lambda0."""
    n = 1 + min(int(rng.uniform(0.0, 6.0)), 5)
    lengths, spacings, radii = [], [], []
    for i in range(n):
        L = rng.uniform(0.3, 0.55) * lambda0
        if i == 0:
            arm = max(2, round_half_away((max(L, 5.0) - 1.0) * 0.5))
            lengths.append(2 * arm + 1)
            spacings.append(0)
        else:
            lengths.append(max(1, round_half_away(L)))
            spacings.append(max(1, round_half_away(rng.uniform(0.1, 0.4) * lambda0)))
        radii.append(float(f32(rng.log_uniform(0.001, 0.01) * lambda0)))
    return ArrayGeometry(lengths, spacings, radii)


def array_scene(geom: ArrayGeometry, size, lambda0: float = ARRAY_LAMBDA0, pml: int = ARRAY_PML,
                sweep_points: int = 64, pulse_width: float = 0.75) -> Scene:
    """Vacuum grid of `size` with the wires of `geom` centred in x/y/z; element 0 is a thin-wire
    dipole behind a 50 ohm resistive feed (Gaussian pulse at lambda0), the others are passive
    wire{} elements; a port sweep over 0.5-1.5 f0."""
    nx, ny, nz = (int(v) for v in size)
    sc = Scene(size=(nx, ny, nz), pml=pml, courant=MAX_COURANT, cell_size=1.0e-3, wavelength=lambda0)
    xs = [0]
    for s in geom.spacings[1:]:
        xs.append(xs[-1] + int(s))
    x0 = nx // 2 - geom.span() // 2
    cy, cz = ny // 2, nz // 2
    sc.dipole_thin((x0 + xs[0], cy, cz + 0.5), axis=2, length=float(geom.lengths[0]), radius=geom.radii[0],
                   impedance=50.0, amplitude=1.0, wavelength=lambda0, waveform="pulse", pulse_width=pulse_width)
    for i in range(1, geom.elements):
        n = int(geom.lengths[i])
        z0 = cz - n // 2
        sc.wire((x0 + xs[i], cy, z0), (x0 + xs[i], cy, z0 + n), radius=geom.radii[i])
    sc.set_sweep((lambda0 / 1.5, lambda0 / 0.5), points=sweep_points)
    sc.emit_thin_wires()
    return sc


def draw_array_design(index: int, seed: int = DEFAULT_SEED, lambda0: float = ARRAY_LAMBDA0):
    """Design `index` of the batched generator: its own xoshiro256** stream (seed * 1000003 +
    index), the geometry first, then the grid size over multiples of 8 in [48,128] x [48,128] x
    [96,160] (x at least the array span plus both absorbers and 8 cells)."""
    rng = Xoshiro256ss(seed * 1000003 + index)
    geom = draw_array_geometry(rng, lambda0)
    lo_x = max(48, -(-(geom.span() + 2 * ARRAY_PML + 8) // 8) * 8)

    def pick(lo, hi):
        k = (hi - lo) // 8 + 1
        return lo + 8 * min(int(rng.uniform(0.0, float(k))), k - 1)

    return geom, (pick(lo_x, 128), pick(48, 128), pick(96, 160))


def draw_array_designs(start: int, count: int, seed: int = DEFAULT_SEED, lambda0: float = ARRAY_LAMBDA0):
    """draw_array_design(i) for i in start .. start+count-1, with the xoshiro256** streams of all
    designs stepped together in numpy uint64 lanes: the same draws, the same double arithmetic
    (the exponentials through math.exp, as the scalar path), the same designs."""
    lane = np.arange(start, start + count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = np.uint64((seed * 1000003) & _M64) + lane
        st = []
        for _ in range(4):
            x = x + np.uint64(0x9E3779B97F4A7C15)
            z = x
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            st.append(z ^ (z >> np.uint64(31)))
        s0, s1, s2, s3 = st

        def rotl(v, k):
            return (v << np.uint64(k)) | (v >> np.uint64(64 - k))

        draws = []  # 3 n + 3 <= 21 draws per design, as uniform [0, 1) doubles
        for _ in range(3 * ARRAY_MAX_ELEMENTS + 3):
            r = rotl(s1 * np.uint64(5), 7) * np.uint64(9)
            t = s1 << np.uint64(17)
            s2 = s2 ^ s0
            s3 = s3 ^ s1
            s1 = s1 ^ s2
            s0 = s0 ^ s3
            s2 = s2 ^ t
            s3 = rotl(s3, 45)
            draws.append((r >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53)))
    u = np.stack(draws, axis=1)

    def uni(col, lo, hi):
        return lo + (hi - lo) * col

    # draw columns: 0 element count; element 0: length 1, radius 2; element i >= 1: length 3i,
    # spacing 3i + 1, radius 3i + 2; then the three grid-size picks at 3n, 3n + 1, 3n + 2
    m = ARRAY_MAX_ELEMENTS
    n = 1 + np.minimum(np.floor(uni(u[:, 0], 0.0, 6.0)).astype(np.int64), 5)
    present = np.arange(m)[None, :] < n[:, None]
    lcol = np.array([1] + [3 * i for i in range(1, m)])
    rcol = np.array([2] + [3 * i + 2 for i in range(1, m)])
    L = uni(u[:, lcol], 0.3, 0.55) * lambda0
    # round_half_away of a positive value: floor(x + 0.5)
    lengths = np.maximum(1, np.floor(L + 0.5)).astype(np.int64)
    arm0 = np.maximum(2, np.floor((np.maximum(L[:, 0], 5.0) - 1.0) * 0.5 + 0.5)).astype(np.int64)
    lengths[:, 0] = 2 * arm0 + 1
    spacings = np.zeros((count, m), np.int64)
    spacings[:, 1:] = np.maximum(1, np.floor(uni(u[:, [3 * i + 1 for i in range(1, m)]], 0.1, 0.4) * lambda0 + 0.5))
    spacings[~present] = 0
    ru = uni(u[:, rcol], math.log(0.001), math.log(0.01))
    radii = np.zeros((count, m), np.float64)
    sel = np.nonzero(present)
    # the exponential through math.exp, as the scalar generator takes it
    radii[sel] = np.array([math.exp(v) for v in ru[sel].tolist()]) * lambda0
    radii = radii.astype(np.float32)
    span = spacings.sum(axis=1)
    lo_x = np.maximum(48, ((span + 2 * ARRAY_PML + 8 + 7) // 8) * 8)
    picks = []
    for q, (lo, hi) in enumerate(((lo_x, 128), (48, 128), (96, 160))):
        kk = (hi - lo) // 8 + 1
        col = np.take_along_axis(u, (3 * n + q)[:, None], axis=1)[:, 0]
        picks.append(lo + 8 * np.minimum(np.floor(kk * col).astype(np.int64), kk - 1))
    geoms, sizes = [], []
    nl, ll, sl, rl = n.tolist(), lengths.tolist(), spacings.tolist(), radii.tolist()
    p0, p1, p2 = (p.tolist() for p in picks)
    for d in range(count):
        nd = nl[d]
        geoms.append(ArrayGeometry(ll[d][:nd], sl[d][:nd], rl[d][:nd]))
        sizes.append((p0[d], p1[d], p2[d]))
    return geoms, sizes


def draw_array_scene(index: int, seed: int = DEFAULT_SEED, lambda0: float = ARRAY_LAMBDA0):
    """(geometry, Scene) of design `index` (draw_array_design)."""
    geom, size = draw_array_design(index, seed, lambda0)
    return geom, array_scene(geom, size, lambda0)


def radar_dipole_array(lam: float = 20.0, target: str = "sphere", range_wl: float = 4.0, az_deg: float = 10.0,
                       el_deg: float = 5.0) -> Scene:
    """A compact radar scene in the manner of scenes/16_array_radar.lua: a 2 x 4 array of z thin-wire
    dipoles (0.45 lambda, 0.5 lambda apart in y, 0.7 lambda in z) looking down +x, the central 2 x 2
    transmitting a pulse, every element receiving, and a metal target (sphere of radius
    0.75 lambda, or a 2 x 2 lambda plate facing the array) `range_wl` wavelengths out."""
    from .radar import RadarSetup
    sc = Scene(size=(200, 112, 112), pml=10, courant=MAX_COURANT, cell_size=1.0e-3, wavelength=lam)
    x0, yc, zc = 40, 56, 56
    for row in (-1, 1):
        for col in range(4):
            y = yc + int(round((col - 1.5) * 10))
            tx = col in (1, 2)
            sc.dipole_thin((x0, y, zc + 7 * row), axis=2, length=9.0, radius=0.1, impedance=50.0,
                           amplitude=1.0 if tx else 0.0, wavelength=lam, waveform="pulse", pulse_width=1.0,
                           name=f"p[{row},{col}]")
    sc.emit_thin_wires()
    r = range_wl * lam
    ra, re = math.radians(az_deg), math.radians(el_deg)
    c = (x0 + r * math.cos(ra) * math.cos(re), yc - r * math.sin(ra) * math.cos(re), zc + r * math.sin(re))
    if target == "sphere":
        sc.sphere(c, 0.75 * lam, conductivity=5.8e7, target=True, name="sphere")
    elif target == "plate":
        sc.box((c[0] - 1, c[1] - lam, c[2] - lam), (c[0] + 1, c[1] + lam, c[2] + lam), conductivity=5.8e7,
               target=True, name="plate")
    else:
        raise ValueError(f"unknown radar target '{target}'")
    sc.radar = RadarSetup(fov_az_deg=30.0, fov_el_deg=30.0, boresight=(1.0, 0.0, 0.0))
    return sc


def crpa_crossed_dipoles(lam: float = 20.0, blind: bool = False) -> Scene:
    """A CRPA scene in the manner of scenes/17_crpa_gps.lua: a 2 x 2 array (0.5 lambda apart) of
    crossed thin-wire dipole pairs (X and Y probes, one cell apart in z so the wires never touch)
    under an RHCP satellite near the zenith and two linearly polarized jammers near the horizon."""
    from .crpa import CrpaElement, CrpaSetup
    sc = Scene(size=(120, 120, 100), pml=10, courant=MAX_COURANT, cell_size=1.0e-3, wavelength=lam)
    elements = []
    for i, (cx, cy) in enumerate([(55, 55), (65, 55), (55, 65), (65, 65)]):
        sc.dipole_thin((cx, cy, 30), axis=0, length=9.0, radius=0.1, impedance=50.0, amplitude=1.0, wavelength=lam,
                       waveform="cw", ramp=3.0, name=f"x{i}")
        sc.dipole_thin((cx, cy, 31), axis=1, length=9.0, radius=0.1, impedance=50.0, amplitude=1.0, wavelength=lam,
                       waveform="cw", ramp=3.0, name=f"y{i}")
        elements.append(CrpaElement(f"x{i}", f"y{i}", (cx + 0.5, cy + 0.5, 30.5)))
    sc.emit_thin_wires()
    sc.emitter((62.0, 58.0, 84.0), name="sat", polarization="rhcp", toward=(60.0, 60.0, 30.5), wavelength=lam,
               waveform="cw", ramp=3.0)
    sc.emitter((108.0, 60.0, 38.0), name="jam east", polarization=(0.0, 1.0, 0.3), wavelength=lam, waveform="cw",
               ramp=3.0)
    sc.emitter((14.0, 96.0, 42.0), name="jam north-west", polarization=(1.0, 1.0, 0.0), wavelength=lam,
               waveform="cw", ramp=3.0)
    sc.crpa = CrpaSetup(satellite="sat", elements=elements, blind=blind, pass_periods=20.0, snapshot_periods=16.0,
                        jnr_db=40.0, snr_db=-20.0, noise_seed=1)
    return sc
