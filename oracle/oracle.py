"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so) — TEST INFRASTRUCTURE ONLY.

Only tests, the smoke check and the benchmark's CPU-baseline leg may import this module. It
drives the scalar C restatement of EM_Lab's shaders in EM_Lab's own texel layout (RGBA per
cell, x fastest).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
_lib = None


class OrGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dt", C.c_float), ("pml", C.c_int),
                ("sigma_max", C.c_float), ("grading", C.c_float)]


class OrSource(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("half_extent", C.c_float * 3), ("softness", C.c_float * 3),
                ("polarization", C.c_float * 3), ("amplitude", C.c_float), ("wavelength", C.c_float),
                ("phase", C.c_float), ("waveform", C.c_int), ("ramp_periods", C.c_float),
                ("repeat_periods", C.c_float), ("delay_periods", C.c_float), ("enabled", C.c_int)]


class OrPort(C.Structure):
    _fields_ = [("texel", C.c_int * 3), ("axis", C.c_int), ("edges", C.c_int), ("sign", C.c_int)]


EDGE_DTYPE = np.dtype([("texel", "<i4", 3), ("axis", "<i4"), ("sigma", "<f4")])
WIRE_DTYPE = np.dtype([("texel", "<i4", 3), ("component", "<i4"), ("term", "<i4"), ("factor", "<f4")])

_P = C.c_void_p
_F = C.c_float
_D = C.c_double
_I = C.c_int
class OrRenderArgs(C.Structure):
    _fields_ = [("inv_vp", C.c_float * 16), ("cam", C.c_float * 3), ("vol_min", C.c_float * 3),
                ("vol_size", C.c_float * 3), ("clip_min", C.c_float * 3), ("clip_max", C.c_float * 3),
                ("n", C.c_int * 3), ("step_voxels", C.c_float), ("gain_e", C.c_float), ("gain_b", C.c_float),
                ("density", C.c_float), ("emission", C.c_float), ("material_opacity", C.c_float),
                ("exposure", C.c_float), ("mode", C.c_int), ("signed_axis", C.c_int), ("width", C.c_int),
                ("height", C.c_int)]


_SIG = {
    "or_exp": (_F, [_F]),
    "or_gradient_noise": (_F, [_F, _F]),
    "or_texture": (_F, [C.c_void_p, _I, _I, _I, _I, _F, _F, _F]),
    "or_box_hit": (None, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(_F), C.POINTER(_F)]),
    "or_render_pixel": (None, [C.POINTER(OrRenderArgs), C.c_void_p, C.c_void_p, _I, _I, C.c_void_p]),
    "or_render_volume": (None, [C.POINTER(OrRenderArgs), C.c_void_p, C.c_void_p, C.c_void_p]),
    "or_f32_to_f16": (C.c_uint16, [_F]),
    "or_f16_to_f32": (_F, [C.c_uint16]),
    "or_pack_half2": (C.c_uint32, [_F, _F]),
    "or_configure": (None, [_I, _I, _I, _F, _I, _F, _F, C.POINTER(OrGrid)]),
    "or_pml_profile": (_F, [C.POINTER(OrGrid), _F]),
    "or_pml_sigma": (_F, [C.POINTER(OrGrid), _F, _F, _F]),
    "or_in_pml": (_I, [C.POINTER(OrGrid), _I, _I, _I]),
    "or_phi": (_F, [_F]),
    "or_upload_materials": (_I, [C.POINTER(OrGrid), _P, _P, _P, C.c_size_t, _P, C.c_size_t, _P, _P, _P, _P]),
    "or_set_sources": (_I, [C.POINTER(OrGrid), _P, _I, _P, _P, _P, C.POINTER(_I), _P, C.POINTER(_I)]),
    "or_signal": (_D, [C.POINTER(OrSource), _D]),
    "or_update_B": (None, [C.POINTER(OrGrid), _P, _P, _P, _I]),
    "or_update_E": (None, [C.POINTER(OrGrid), _P, _P, _P]),
    "or_apply_sources": (None, [C.POINTER(OrGrid), _P, _P, _P, _P, _I, _P, _I, _P]),
    "or_clear_fields": (None, [C.POINTER(OrGrid), _P, _P]),
    "or_step": (None, [C.POINTER(OrGrid), _P, _P, _P, _P, _I, _P, _I, _I, C.POINTER(C.c_uint64)]),
    "or_step_storage": (None, [C.POINTER(OrGrid), _P, _P, _P, _P, _I, _P, _I, _I, C.POINTER(C.c_uint64), _I, _P]),
    "or_quantize_fields": (None, [C.POINTER(OrGrid), _P, _P, _I, _P]),
    "or_round_f16": (C.c_float, [C.c_float]),
    "or_round_bf16": (C.c_float, [C.c_float]),
    "or_port_accumulate": (None, [C.POINTER(OrGrid), _P, _P, _P, _I, _P, _F, _I, _P, _P]),
    "or_monitor_constants": (None, [_D, _F, _F, C.c_uint64, _P, C.POINTER(_F)]),
    "or_line_accumulate": (None, [C.POINTER(OrGrid), _P, _P, _P, _I, _P, _P, _P, _F, _P, _P]),
    "or_ntff_point": (None, [_P, _P, _P, _I, _P, _P, C.POINTER(_F)]),
    "or_ntff_face_offsets": (_I, [_P, _P, _P]),
    "or_ntff_accumulate": (None, [C.POINTER(OrGrid), _P, _P, _P, _P, _P, _I, _P, _F, _P]),
    "or_ntff_eval_f32": (None, [_P, _P, _P, _I, _P, _P, _I, _F, _P, _P]),
    "or_ntff_eval_f64": (None, [_P, _P, _P, _I, _P, _P, _I, _D, _P, _P, C.POINTER(_D)]),
    "or_build_directions": (_I, [_P, _P, _I, _P, _P]),
    "or_ntff_samples": (None, [C.POINTER(OrGrid), _P, _P, _P, _P, _P, _I, _P]),
    "or_cpml_coefficients": (None, [C.POINTER(OrGrid), _D, _D, _D, _P]),
    "or_cpml_b": (None, [C.POINTER(OrGrid), _P, _P, _P, _P]),
    "or_cpml_e": (None, [C.POINTER(OrGrid), _P, _P, _P, _P, _P]),
    "or_step_cpml": (None, [C.POINTER(OrGrid), _P, _P, _P, _P, _I, _P, _I, _I, C.POINTER(C.c_uint64), _P, _P]),
    "or_cpml_psi_floats": (_I, [C.POINTER(OrGrid)]),
    "or_e_at_node": (None, [C.POINTER(OrGrid), _P, _I, _I, _I, _P]),
    "or_b_at_node": (None, [C.POINTER(OrGrid), _P, _I, _I, _I, _P]),
    "or_plane_flux": (_D, [C.POINTER(OrGrid), _P, _P, _I, _I, _I, _I, _I, _I]),
    "or_box_flux": (_D, [C.POINTER(OrGrid), _P, _P, _P, _P]),
    "or_probe": (None, [C.POINTER(OrGrid), _P, _P, _P, _P]),
    "or_sweep": (_D, [_P, C.c_size_t, _I, _I, _D, _D, _D, _I, _D, _P, _P]),
}


def build() -> Path:
    r = subprocess.run(["make", "-s", "-C", str(HERE)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        l = C.CDLL(str(LIB))
        for name, (res, args) in _SIG.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def grid(nx, ny, nz, courant, pml, sigma_max=1.6, grading=3.0) -> OrGrid:
    g = OrGrid()
    lib().or_configure(nx, ny, nz, courant, pml, sigma_max, grading, C.byref(g))
    return g


def source_c(s) -> OrSource:
    f3 = C.c_float * 3
    return OrSource(f3(*s.center), f3(*s.half_extent), f3(*s.softness), f3(*s.polarization), s.amplitude,
                    s.wavelength, s.phase, int(s.waveform), s.ramp_periods, s.repeat_periods, s.delay_periods,
                    1 if s.enabled else 0)


def edge_array(edges):
    a = np.zeros(len(edges), EDGE_DTYPE)
    for i, (t, ax, s) in enumerate(edges):
        a[i] = (t, ax, s)
    return a


def wire_array(wires):
    a = np.zeros(len(wires), WIRE_DTYPE)
    for i, (t, comp, term, f) in enumerate(wires):
        a[i] = (t, comp, term, f)
    return a


class OracleSim:
    """A whole simulation on the CPU oracle: materials, sources, steps and per-step monitors,
    in the order of SimulationGrid::step (src/SimulationGrid.cpp:368-433)."""

    def __init__(self, size, courant, pml, sigma_max=1.6, grading=3.0):
        L = lib()
        self.g = grid(size[0], size[1], size[2], courant, pml, sigma_max, grading)
        n = self.g.nx * self.g.ny * self.g.nz
        self.n = n
        self.E = np.zeros((n, 4), np.float32)
        self.B = np.zeros((n, 4), np.float32)
        self.B[:, 3] = 1.0
        self.edge = np.zeros((n, 4), np.uint32)
        self.edge[:, 0] = L.or_pack_half2(0.0, 0.0)
        self.edge[:, 1] = L.or_pack_half2(0.0, 1.0)
        self.edge[:, 2] = L.or_pack_half2(1.0, 1.0)
        self.wire = np.zeros((1, 4), np.uint32)
        self.thin = 0
        self.step_count = 0
        self.sources = []
        self.src_c = None
        self.ports = None
        self.ntff = None
        self.storage = 0
        self.scale_log2 = np.zeros(6, np.int32)

    def set_storage(self, storage, scale_log2=None):
        """Emulates the library's 16-bit field storage (1: binary16, 2: bfloat16): clears the
        fields and rewinds time, as emfd_grid_set_storage."""
        self.storage = int(storage)
        self.scale_log2 = np.zeros(6, np.int32) if scale_log2 is None else np.asarray(scale_log2, np.int32).copy()
        self.E[:, :3] = 0.0
        self.B[:, :3] = 0.0
        self.step_count = 0

    def set_cpml(self, sigma_max, grading, alpha_max):
        """The library's opt-in convolutional PML in place of the graded absorber (same slabs):
        the graded conductivity is switched off and the auxiliary fields start at 0; clears the
        fields and rewinds time, as emfd_grid_set_absorber."""
        L = lib()
        self.g.sigma_max = 0.0
        stride = max(self.g.nx, self.g.ny, self.g.nz)
        self.cpml_coef = np.zeros(3 * 2 * stride * 2, np.float32)
        # the library takes the parameters as floats
        f32 = lambda v: float(np.float32(v))  # noqa: E731
        L.or_cpml_coefficients(C.byref(self.g), f32(sigma_max), f32(grading), f32(alpha_max), ptr(self.cpml_coef))
        self.cpml_psi = np.zeros(L.or_cpml_psi_floats(C.byref(self.g)), np.float32)
        self.E[:, :3] = 0.0
        self.B[:, :3] = 0.0
        self.step_count = 0

    @property
    def size(self):
        return (self.g.nx, self.g.ny, self.g.nz)

    def upload_materials(self, sigma, eps, edges=(), wires=()):
        L = lib()
        s = np.ascontiguousarray(sigma, np.float32).ravel()
        e = np.ascontiguousarray(eps, np.float32).ravel()
        ea = edge_array(list(edges))
        wa = wire_array(list(wires))
        self.wire = np.zeros((self.n if len(wa) else 1, 4), np.uint32)
        self.thin = L.or_upload_materials(C.byref(self.g), ptr(s), ptr(e), ptr(ea) if len(ea) else None, len(ea),
                                          ptr(wa) if len(wa) else None, len(wa), ptr(self.E), ptr(self.B),
                                          ptr(self.edge), ptr(self.wire))
        self.step_count = 0

    def set_sources(self, sources):
        self.sources = list(sources)
        self.src_c = (OrSource * max(len(sources), 1))(*[source_c(s) for s in sources])

    def configure_ports(self, ports, averaging, record=False, capacity=0):
        arr = (OrPort * max(len(ports), 1))()
        for i, p in enumerate(ports):
            arr[i] = OrPort((C.c_int * 3)(*p.texel), p.axis, p.edges, 1 if p.sign >= 0 else -1)
        wl = max(ports[0].wavelength, 2.0) if ports else 16.0
        self.ports = dict(arr=arr, n=len(ports), averaging=max(averaging, 0.5), wavelength=np.float32(wl),
                          phasors=np.zeros((max(len(ports), 1), 8), np.float32), record=record,
                          series=np.zeros((capacity, max(len(ports), 1), 2), np.float32), recorded=0,
                          capacity=capacity)

    def configure_ntff(self, box_min, box_max, wavelength, averaging):
        L = lib()
        bmin = np.array(box_min, np.int32)
        bmax = np.array(box_max, np.int32)
        fo = np.zeros(7, np.int32)
        npts = L.or_ntff_face_offsets(ptr(bmin), ptr(bmax), ptr(fo))
        self.ntff = dict(bmin=bmin, bmax=bmax, fo=fo, n=npts, wavelength=np.float32(max(wavelength, 2.0)),
                         averaging=max(averaging, 0.5), phasors=np.zeros((npts, 32), np.float32))
        self.ntff_extra = []

    def add_ntff_wavelength(self, wavelength):
        """A further far-field accumulator on the same box at another wavelength (the library's
        multi-wavelength monitor keeps one per wavelength); returns its phasor array."""
        N = self.ntff
        extra = dict(wavelength=np.float32(max(wavelength, 2.0)), phasors=np.zeros((N["n"], 32), np.float32))
        self.ntff_extra.append(extra)
        return extra["phasors"]

    def configure_ntff_dft(self, wavelengths):
        """FP64 running DFTs of the Huygens-node samples at these wavelengths (the library's opt-in
        far-field DFT): per wavelength (n_points, 3) complex128 sums for E at t = n dt and H at
        (n - 1/2) dt, n being the step count after the step. The DFT is summed here in numpy."""
        N = self.ntff
        self.ntff_dft = dict(wavelengths=[float(w) for w in wavelengths],
                             E=np.zeros((len(wavelengths), N["n"], 3), np.complex128),
                             H=np.zeros((len(wavelengths), N["n"], 3), np.complex128))

    def ntff_samples(self):
        N = self.ntff
        out = np.zeros((N["n"], 6), np.float32)
        lib().or_ntff_samples(C.byref(self.g), ptr(self.E), ptr(self.B), ptr(N["bmin"]), ptr(N["bmax"]), ptr(N["fo"]),
                              N["n"], ptr(out))
        return out

    def step(self, count):
        L = lib()
        if count <= 0:
            return
        # one call per step keeps the per-step monitors in the reference's order
        for _ in range(count):
            sc = C.c_uint64(self.step_count)
            if getattr(self, "cpml_psi", None) is not None:
                L.or_step_cpml(C.byref(self.g), ptr(self.E), ptr(self.B), ptr(self.edge), ptr(self.wire), self.thin,
                               C.cast(self.src_c, C.c_void_p) if self.src_c else None, len(self.sources), 1,
                               C.byref(sc), ptr(self.cpml_coef), ptr(self.cpml_psi))
            else:
                L.or_step_storage(C.byref(self.g), ptr(self.E), ptr(self.B), ptr(self.edge), ptr(self.wire),
                                  self.thin, C.cast(self.src_c, C.c_void_p) if self.src_c else None,
                                  len(self.sources), 1, C.byref(sc), self.storage, ptr(self.scale_log2))
            self.step_count = sc.value
            self._monitors()

    def _monitors(self):
        L = lib()
        if self.ports and self.ports["n"]:
            P = self.ports
            rot = np.zeros(4, np.float32)
            blend = C.c_float()
            L.or_monitor_constants(float(P["wavelength"]), P["averaging"], self.g.dt, self.step_count, ptr(rot),
                                   C.byref(blend))
            rec = -1
            if P["record"] and P["recorded"] < P["capacity"]:
                rec = P["recorded"]
                P["recorded"] += 1
            L.or_port_accumulate(C.byref(self.g), ptr(self.E), ptr(self.B), C.cast(P["arr"], C.c_void_p), P["n"],
                                 ptr(rot), blend.value, rec, ptr(P["phasors"]), ptr(P["series"]))
        if self.ntff:
            N = self.ntff
            rot = np.zeros(4, np.float32)
            blend = C.c_float()
            L.or_monitor_constants(float(N["wavelength"]), N["averaging"], self.g.dt, self.step_count, ptr(rot),
                                   C.byref(blend))
            L.or_ntff_accumulate(C.byref(self.g), ptr(self.E), ptr(self.B), ptr(N["bmin"]), ptr(N["bmax"]),
                                 ptr(N["fo"]), N["n"], ptr(rot), blend.value, ptr(N["phasors"]))
            for X in getattr(self, "ntff_extra", []):
                L.or_monitor_constants(float(X["wavelength"]), N["averaging"], self.g.dt, self.step_count, ptr(rot),
                                       C.byref(blend))
                L.or_ntff_accumulate(C.byref(self.g), ptr(self.E), ptr(self.B), ptr(N["bmin"]), ptr(N["bmax"]),
                                     ptr(N["fo"]), N["n"], ptr(rot), blend.value, ptr(X["phasors"]))
        D = getattr(self, "ntff_dft", None)
        if D is not None:
            smp = self.ntff_samples().astype(np.float64)
            dt = float(np.float32(self.g.dt))
            tE = float(self.step_count) * dt
            tH = tE - 0.5 * dt
            for q, wl in enumerate(D["wavelengths"]):
                w = 2.0 * np.pi / wl
                D["E"][q] += smp[:, :3] * complex(np.cos(w * tE), -np.sin(w * tE))
                D["H"][q] += smp[:, 3:] * complex(np.cos(w * tH), -np.sin(w * tH))
        if getattr(self, "lines", None):
            Q = self.lines
            rot = np.zeros(4, np.float32)
            scratch = np.zeros(4, np.float32)
            blend = C.c_float()
            unused = C.c_float()
            # src/Coax.cpp:177-182: the rotation at max(wavelength, 2), the blend at the wavelength
            L.or_monitor_constants(max(Q["wavelength"], 2.0), Q["averaging"], self.g.dt, self.step_count, ptr(rot),
                                   C.byref(unused))
            L.or_monitor_constants(float(Q["wavelength"]), Q["averaging"], self.g.dt, self.step_count, ptr(scratch),
                                   C.byref(blend))
            L.or_line_accumulate(C.byref(self.g), ptr(self.E), ptr(self.B), ptr(Q["quantities"]), Q["n"],
                                 ptr(Q["terms"]), ptr(Q["weights"]), ptr(rot), blend.value, ptr(Q["phasors"]),
                                 ptr(Q["now"]))

    def configure_lines(self, quantities, terms, weights, wavelength, averaging):
        """CoaxMonitor::configure's device state (src/Coax.cpp:95-188): quantities (n, 4) int32,
        terms (m, 4) int32, weights (m,) float32; phasors and instantaneous values start at 0."""
        q = np.ascontiguousarray(quantities, np.int32).reshape(-1, 4)
        self.lines = dict(quantities=q, n=q.shape[0], terms=np.ascontiguousarray(terms, np.int32).reshape(-1, 4),
                          weights=np.ascontiguousarray(weights, np.float32), wavelength=float(wavelength),
                          averaging=max(float(averaging), 0.5), phasors=np.zeros((q.shape[0], 4), np.float32),
                          now=np.zeros(q.shape[0], np.float32))

    def ntff_eval_f32(self, dirs, k=None, extra=None):
        """Pattern |N_perp - r x L|^2 and power of the primary accumulator, or of the extra
        wavelength `extra` (index into add_ntff_wavelength's list)."""
        L = lib()
        N = self.ntff
        src = N if extra is None else self.ntff_extra[extra]
        d = np.ascontiguousarray(dirs, np.float32)
        out = np.zeros(2 * d.shape[0], np.float32)
        kk = np.float32(2.0 * np.pi / float(src["wavelength"])) if k is None else np.float32(k)
        center = np.array([0.5 * (N["bmin"][a] + N["bmax"][a]) for a in range(3)], np.float32)
        L.or_ntff_eval_f32(ptr(N["bmin"]), ptr(N["bmax"]), ptr(N["fo"]), N["n"], ptr(src["phasors"]), ptr(d), d.shape[0],
                           float(kk), ptr(center), ptr(out))
        return out[0::2], float(out[1])

    def plane_flux(self, axis, index, rmin, rmax):
        return lib().or_plane_flux(C.byref(self.g), ptr(self.E), ptr(self.B), axis, index, rmin[0], rmin[1], rmax[0],
                                   rmax[1])

    def box_flux(self, bmin, bmax):
        a = np.array(bmin, np.int32)
        b = np.array(bmax, np.int32)
        return lib().or_box_flux(C.byref(self.g), ptr(self.E), ptr(self.B), ptr(a), ptr(b))

    def probe(self, pos):
        out = np.zeros(11, np.float32)
        p = np.array(pos, np.float32)
        lib().or_probe(C.byref(self.g), ptr(self.E), ptr(self.B), ptr(p), ptr(out))
        return out


def render_args(size, settings, inv_vp, cam, width, height) -> OrRenderArgs:
    """The raymarch's inputs after the viewer's set-up (src/VolumeRenderer.cpp:44-54,89-112):
    world box (longest axis 2 units, centred), cutaway and step clamps. settings: the fields of
    RenderSettings (mode, signed_axis, gain_e, ...)."""
    r = OrRenderArgs()
    n = [int(v) for v in size]
    ws = np.float32(2.0) / np.float32(max(max(n), 1))
    for a in range(3):
        r.n[a] = n[a]
        r.cam[a] = float(cam[a])
        vmin = np.float32(-0.5) * np.float32(n[a]) * ws
        vsz = np.float32(n[a]) * ws
        lo = min(max(min(settings["clip_min"][a], settings["clip_max"][a]), 0.0), 1.0)
        hi = min(max(max(settings["clip_min"][a], settings["clip_max"][a]), 0.0), 1.0)
        r.vol_min[a], r.vol_size[a] = float(vmin), float(vsz)
        r.clip_min[a] = float(vmin + np.float32(lo) * vsz)
        r.clip_max[a] = float(vmin + np.float32(hi) * vsz)
    r.inv_vp[:] = [float(v) for v in inv_vp]
    r.step_voxels = min(max(settings["step_voxels"], 0.1), 4.0)
    for k in ("gain_e", "gain_b", "density", "emission", "material_opacity", "exposure"):
        setattr(r, k, settings[k])
    r.mode = int(settings["mode"])
    r.signed_axis = min(max(int(settings["signed_axis"]), 0), 2)
    r.width, r.height = int(width), int(height)
    return r


def render_volume(args: OrRenderArgs, E: np.ndarray, B: np.ndarray) -> np.ndarray:
    """(height, width, 4) float image, bottom row first; E, B: RGBA texels (n, 4) float32."""
    E = np.ascontiguousarray(E, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    out = np.zeros((args.height, args.width, 4), np.float32)
    lib().or_render_volume(C.byref(args), ptr(E), ptr(B), ptr(out))
    return out
