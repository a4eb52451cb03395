"""Python host binding of libemfdtd.so (ctypes over the C ABI in include/emfdtd.h).

The classes keep EM_Lab's names and call order (src/SimulationGrid.hpp:155-216,
src/Ports.hpp:84-140, src/FarField.hpp:75-133, src/FieldAnalyzer.hpp:73-130). Every call
goes to the HIP library; there is no CPU fallback. If the library is missing or no GPU is
present, construction raises.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libemfdtd.so"
_lib = None

EMFD_OK = 0
EMFD_ERR_INVALID = 1
EMFD_ERR_HIP = 2
EMFD_ERR_STATE = 3
LAYOUT_TEXEL = 0
LAYOUT_SOA = 1
KERNEL_AUTO = 0
KERNEL_TWO_PASS = 2
KERNEL_TWO_PASS4 = 5
KERNEL_MARCH = 8
ABSORBER_GRADED = 0
ABSORBER_CPML = 1
STORAGE_F32 = 0
STORAGE_F16 = 1
STORAGE_BF16 = 2
STORAGE_F64 = 3


class EmfdError(RuntimeError):
    """Raised for any non-zero status of the C ABI (message from emfd_last_error)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"emfdtd error {code}: {msg}")
        self.code = code


# ---------------------------------------------------------------------------------------------
# C structures (must match include/emfdtd.h)
# ---------------------------------------------------------------------------------------------
class GridConfigC(C.Structure):
    _fields_ = [("size", C.c_int * 3), ("courant", C.c_float), ("pml_thickness", C.c_int),
                ("pml_sigma_max", C.c_float), ("pml_grading", C.c_float)]


class SourceDescC(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("half_extent", C.c_float * 3), ("softness", C.c_float * 3),
                ("polarization", C.c_float * 3), ("amplitude", C.c_float), ("wavelength", C.c_float),
                ("phase", C.c_float), ("waveform", C.c_int), ("ramp_periods", C.c_float),
                ("repeat_periods", C.c_float), ("delay_periods", C.c_float), ("enabled", C.c_int)]


class PortSetupC(C.Structure):
    _fields_ = [("texel", C.c_int * 3), ("axis", C.c_int), ("sign", C.c_float), ("edges", C.c_int),
                ("wavelength", C.c_float), ("voltage_feed", C.c_int), ("impedance_ohm", C.c_float),
                ("drive", C.c_float)]


class SweepSetupC(C.Structure):
    _fields_ = [("wavelength_min", C.c_float), ("wavelength_max", C.c_float), ("points", C.c_int)]


class PortReadingC(C.Structure):
    _fields_ = [("valid", C.c_int)] + [(n, C.c_double) for n in (
        "v_re", "v_im", "i_re", "i_im", "z_re", "z_im", "gamma_re", "gamma_im", "vswr", "return_loss_db",
        "accepted_power", "available_power")]


class SweepPointC(C.Structure):
    _fields_ = [("wavelength", C.c_double), ("z_re", C.c_double), ("z_im", C.c_double), ("s11_db", C.c_double),
                ("vswr", C.c_double)]


class PortSweepC(C.Structure):
    _fields_ = [("best", C.c_int64), ("best_wavelength", C.c_double), ("has_band", C.c_int),
                ("band_low_wavelength", C.c_double), ("band_high_wavelength", C.c_double)]


class FarFieldSetupC(C.Structure):
    _fields_ = [("box_min", C.c_int * 3), ("box_max", C.c_int * 3), ("wavelength", C.c_float),
                ("averaging_periods", C.c_float), ("map_axis", C.c_float * 3)]


class FluxSurfaceC(C.Structure):
    _fields_ = [("kind", C.c_int), ("axis", C.c_int), ("index", C.c_int), ("rect_min", C.c_int * 2),
                ("rect_max", C.c_int * 2), ("box_min", C.c_int * 3), ("box_max", C.c_int * 3)]


class ProbeSampleC(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("E", C.c_float * 3), ("B", C.c_float * 3), ("S", C.c_float * 3),
                ("sigma", C.c_float), ("epsilon_r", C.c_float)]


class RadarImageSetupC(C.Structure):
    _fields_ = [("tx", C.c_float * 3), ("axes", C.c_float * 9), ("interior_min", C.c_float * 3),
                ("interior_max", C.c_float * 3), ("wavelength", C.c_float), ("dt", C.c_double),
                ("sample_dt", C.c_double), ("r0", C.c_float), ("dr", C.c_float), ("nr", C.c_int),
                ("az0", C.c_float), ("daz", C.c_float), ("naz", C.c_int), ("el0", C.c_float), ("del_", C.c_float),
                ("nel", C.c_int)]


class FieldViewC(C.Structure):
    _fields_ = [("size", C.c_int * 3), ("storage", C.c_int), ("e", C.c_void_p * 3), ("b", C.c_void_p * 3),
                ("scale", C.c_float * 6), ("sigma", C.c_void_p), ("eps_r", C.c_void_p), ("device", C.c_int),
                ("step_count", C.c_uint64)]


class RenderSettingsC(C.Structure):
    _fields_ = [("mode", C.c_int), ("signed_axis", C.c_int), ("gain_e", C.c_float), ("gain_b", C.c_float),
                ("density", C.c_float), ("emission", C.c_float), ("step_voxels", C.c_float),
                ("material_opacity", C.c_float), ("exposure", C.c_float), ("clip_min", C.c_float * 3),
                ("clip_max", C.c_float * 3)]


class BatchSceneC(C.Structure):
    _fields_ = [("config", GridConfigC), ("sigma", C.c_void_p), ("eps_r", C.c_void_p), ("n_cells", C.c_int64),
                ("edges", C.c_void_p), ("n_edges", C.c_int64), ("wires", C.c_void_p), ("n_wires", C.c_int64),
                ("sources", C.c_void_p), ("n_sources", C.c_int), ("ports", C.c_void_p), ("n_ports", C.c_int),
                ("averaging_periods", C.c_float), ("sweep", C.c_void_p)]


class BatchLaunchC(C.Structure):
    _fields_ = [("quads_per_thread", C.c_int), ("persistent_blocks", C.c_int), ("xcd_slices", C.c_int)]


EDGE_DTYPE = np.dtype([("texel", "<i4", 3), ("axis", "<i4"), ("sigma", "<f4")])
WIRE_DTYPE = np.dtype([("texel", "<i4", 3), ("component", "<i4"), ("term", "<i4"), ("factor", "<f4")])

_P = C.c_void_p
_SIGNATURES = {
    "emfd_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "emfd_destroy": (None, [_P]),
    "emfd_last_error": (C.c_char_p, []),
    "emfd_version": (C.c_char_p, []),
    "emfd_grid_configure": (C.c_int, [_P, C.POINTER(GridConfigC), C.POINTER(GridConfigC)]),
    "emfd_grid_get_config": (C.c_int, [_P, C.POINTER(GridConfigC)]),
    "emfd_grid_upload_materials": (C.c_int, [_P, _P, _P, C.c_int64, _P, C.c_int64, _P, C.c_int64]),
    "emfd_grid_set_sources": (C.c_int, [_P, _P, C.c_int, C.POINTER(C.c_int)]),
    "emfd_grid_reset_fields": (C.c_int, [_P]),
    "emfd_grid_step": (C.c_int, [_P, C.c_int]),
    "emfd_grid_synchronize": (C.c_int, [_P]),
    "emfd_grid_state": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.POINTER(C.c_float)]),
    "emfd_grid_download_fields": (C.c_int, [_P, _P, _P, C.c_int]),
    "emfd_grid_upload_fields": (C.c_int, [_P, _P, _P, C.c_int]),
    "emfd_grid_memory_bytes": (C.c_int64, [_P]),
    "emfd_grid_set_kernel": (C.c_int, [_P, C.c_int]),
    "emfd_grid_material_classes": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "emfd_grid_set_absorber": (C.c_int, [_P, C.c_int, C.c_float, C.c_float, C.c_float]),
    "emfd_grid_set_storage": (C.c_int, [_P, C.c_int, _P]),
    "emfd_grid_download_fields_f64": (C.c_int, [_P, _P, _P]),
    "emfd_grid_kernel_timing": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "emfd_grid_kernel_timing_samples": (C.c_int, [_P, _P, C.c_int, C.POINTER(C.c_int)]),
    "emfd_lines_configure": (C.c_int, [_P, _P, C.c_int, _P, _P, C.c_int, C.c_float, C.c_float]),
    "emfd_lines_read": (C.c_int, [_P, _P, _P, C.c_int]),
    "emfd_coax_geometry_factor": (C.c_double, [C.c_int, C.c_int]),
    "emfd_ports_configure": (C.c_int, [_P, _P, C.c_int, C.c_float, _P, C.c_float, C.c_int64]),
    "emfd_ports_disable": (C.c_int, [_P]),
    "emfd_ports_read_phasors": (C.c_int, [_P, _P, C.c_int]),
    "emfd_ports_set_dft": (C.c_int, [_P, _P, C.c_int]),
    "emfd_ports_read_dft": (C.c_int, [_P, _P, C.c_int64]),
    "emfd_ports_evaluate": (C.c_int, [_P, _P, C.c_int]),
    "emfd_ports_recorded_steps": (C.c_int64, [_P]),
    "emfd_ports_read_series": (C.c_int, [_P, _P, C.c_int64]),
    "emfd_ports_evaluate_sweep": (C.c_int, [_P, C.c_double, _P, _P, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                            C.POINTER(C.c_int)]),
    "emfd_ntff_configure": (C.c_int, [_P, C.POINTER(FarFieldSetupC), C.POINTER(FarFieldSetupC), C.POINTER(C.c_int)]),
    "emfd_ntff_disable": (C.c_int, [_P]),
    "emfd_ntff_reset": (C.c_int, [_P]),
    "emfd_ntff_read_phasors": (C.c_int, [_P, _P, C.c_int64]),
    "emfd_ntff_directions": (C.c_int, [_P, _P, C.c_int, C.POINTER(C.c_int), _P, _P]),
    "emfd_ntff_eval": (C.c_int, [_P, _P, C.c_int, _P, C.POINTER(C.c_double)]),
    "emfd_ntff_set_wavelengths": (C.c_int, [_P, _P, C.c_int]),
    "emfd_ntff_eval_at": (C.c_int, [_P, C.c_int, _P, C.c_int, _P, C.POINTER(C.c_double)]),
    "emfd_ntff_read_phasors_at": (C.c_int, [_P, C.c_int, _P, C.c_int64]),
    "emfd_ntff_set_dft": (C.c_int, [_P, C.c_int]),
    "emfd_ntff_read_dft": (C.c_int, [_P, C.c_int, _P, C.c_int64]),
    "emfd_ntff_eval_dft": (C.c_int, [_P, C.c_int, _P, C.c_int, _P, C.POINTER(C.c_double)]),
    "emfd_ntff_evaluate": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_float), _P, _P,
                                     _P]),
    "emfd_flux_measure": (C.c_int, [_P, _P, C.c_int, _P]),
    "emfd_flux_measure_noalloc": (C.c_int, [_P, _P, C.c_int, _P]),
    "emfd_probe_measure": (C.c_int, [_P, _P, C.c_int, _P]),
    "emfd_copy_probe": (C.c_int, [_P, C.c_int64, C.c_int, C.POINTER(C.c_double)]),
    "emfd_stream_probe": (C.c_int, [_P, C.c_int64, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "emfd_batch_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "emfd_batch_destroy": (C.c_int, [_P]),
    "emfd_batch_add_scene": (C.c_int, [_P, C.POINTER(BatchSceneC), C.POINTER(C.c_int), C.POINTER(GridConfigC)]),
    "emfd_batch_finalize": (C.c_int, [_P, C.c_int64]),
    "emfd_batch_info": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_size_t)]),
    "emfd_batch_set_launch": (C.c_int, [_P, C.POINTER(BatchLaunchC)]),
    "emfd_batch_get_launch": (C.c_int, [_P, C.POINTER(BatchLaunchC)]),
    "emfd_batch_autotune": (C.c_int, [_P, C.c_int, C.POINTER(BatchLaunchC), _P, C.c_int, C.POINTER(C.c_int)]),
    "emfd_batch_set_graph": (C.c_int, [_P, C.c_int]),
    "emfd_batch_step": (C.c_int, [_P, C.c_int]),
    "emfd_batch_synchronize": (C.c_int, [_P]),
    "emfd_batch_state": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "emfd_batch_reset": (C.c_int, [_P]),
    "emfd_batch_download_fields": (C.c_int, [_P, C.c_int, _P, _P]),
    "emfd_batch_upload_fields": (C.c_int, [_P, C.c_int, _P, _P]),
    "emfd_batch_read_phasors": (C.c_int, [_P, C.c_int, _P, C.c_int]),
    "emfd_batch_read_series": (C.c_int, [_P, C.c_int, _P, C.c_int64]),
    "emfd_batch_evaluate_sweeps": (C.c_int, [_P]),
    "emfd_batch_sweep_result": (C.c_int, [_P, C.c_int, _P, C.c_int, _P, C.POINTER(C.c_double)]),
    "emfd_batch_kernel_timing": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "emfd_mlp_create": (C.c_int, [C.c_int, _P, C.c_int, C.c_int, C.POINTER(_P)]),
    "emfd_mlp_destroy": (C.c_int, [_P]),
    "emfd_mlp_param_count": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "emfd_mlp_set_params": (C.c_int, [_P, _P, C.c_int64]),
    "emfd_mlp_get_params": (C.c_int, [_P, _P, C.c_int64]),
    "emfd_mlp_forward": (C.c_int, [_P, _P, C.c_int64, _P]),
    "emfd_mlp_gradients": (C.c_int, [_P, _P, _P, C.c_int64, _P, C.POINTER(C.c_double)]),
    "emfd_mlp_set_adam": (C.c_int, [_P, C.c_float, C.c_float, C.c_float, C.c_float]),
    "emfd_mlp_train_step": (C.c_int, [_P, _P, _P, C.c_int64, C.c_float, C.POINTER(C.c_double)]),
    "emfd_mlp_set_dataset": (C.c_int, [_P, _P, _P, C.c_int64]),
    "emfd_mlp_train": (C.c_int, [_P, C.c_int, C.c_int, C.c_float, C.c_float, C.c_uint64, _P, C.c_int]),
    "emfd_mlp_time_forward": (C.c_int, [_P, C.c_int64, C.c_int, C.POINTER(C.c_double)]),
    "emfd_radar_backproject": (C.c_int, [_P, _P, _P, C.c_int64, C.c_int, _P, _P, _P, _P, C.c_int, _P]),
    "emfd_grid_field_view": (C.c_int, [_P, C.POINTER(FieldViewC)]),
    "emfd_render_default_settings": (None, [C.POINTER(RenderSettingsC)]),
    "emfd_render_volume": (C.c_int, [_P, C.POINTER(RenderSettingsC), _P, _P, C.c_int, C.c_int, C.c_int, _P]),
}


def library_path() -> Path:
    return _LIB_PATH


def load_library() -> C.CDLL:
    """Load libemfdtd.so once. Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise RuntimeError(f"{_LIB_PATH} is missing: build it with `python -m em_lab_amd.build`")
    lib = C.CDLL(str(_LIB_PATH))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != EMFD_OK:
        raise EmfdError(rc, _lib.emfd_last_error().decode(errors="replace"))


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------------------------------------
# Plain data mirrors of the reference's structs
# ---------------------------------------------------------------------------------------------
@dataclass
class GridConfig:
    """GridConfig (src/SimulationGrid.hpp:25-31)."""
    size: tuple = (128, 128, 128)
    courant: float = 0.5
    pml_thickness: int = 16
    pml_sigma_max: float = 1.6
    pml_grading: float = 3.0

    def to_c(self) -> GridConfigC:
        return GridConfigC((C.c_int * 3)(*[int(v) for v in self.size]), self.courant, int(self.pml_thickness),
                           self.pml_sigma_max, self.pml_grading)

    @staticmethod
    def from_c(c: GridConfigC) -> "GridConfig":
        return GridConfig(tuple(c.size), c.courant, c.pml_thickness, c.pml_sigma_max, c.pml_grading)


WAVE_CW = 0
WAVE_PULSE = 1


@dataclass
class SourceDesc:
    """SourceDesc (src/SimulationGrid.hpp:113-150); factories as src/SimulationGrid.cpp:58-97."""
    center: tuple = (0.0, 0.0, 0.0)
    half_extent: tuple = (0.0, 0.0, 0.0)
    softness: tuple = (0.75, 0.75, 0.75)
    polarization: tuple = (0.0, 0.0, 1.0)
    amplitude: float = 1.0
    wavelength: float = 16.0
    phase: float = 0.0
    waveform: int = WAVE_CW
    ramp_periods: float = 3.0
    repeat_periods: float = 0.0
    delay_periods: float = 0.0
    enabled: bool = True
    port: int = -1
    emitter: int = -1  # emitter{} index (src/SimulationGrid.hpp:149), host-side only

    def to_c(self) -> SourceDescC:
        f3 = C.c_float * 3
        return SourceDescC(f3(*self.center), f3(*self.half_extent), f3(*self.softness), f3(*self.polarization),
                           self.amplitude, self.wavelength, self.phase, int(self.waveform), self.ramp_periods,
                           self.repeat_periods, self.delay_periods, 1 if self.enabled else 0)


@dataclass
class PortSetup:
    """PortSetup (src/Ports.hpp:30-40)."""
    name: str = "port"
    texel: tuple = (0, 0, 0)
    axis: int = 2
    sign: float = 1.0
    edges: int = 1
    wavelength: float = 16.0
    voltage_feed: bool = True
    impedance_ohm: float = 50.0
    drive: float = 1.0

    def to_c(self) -> PortSetupC:
        return PortSetupC((C.c_int * 3)(*self.texel), self.axis, self.sign, self.edges, self.wavelength,
                          1 if self.voltage_feed else 0, self.impedance_ohm, self.drive)


@dataclass
class SweepSetup:
    """SweepSetup (src/Ports.hpp:42-47)."""
    wavelength_min: float = 12.0
    wavelength_max: float = 20.0
    points: int = 81


@dataclass
class FarFieldSetup:
    """FarFieldSetup (src/FarField.hpp:39-47)."""
    box_min: tuple = (0, 0, 0)
    box_max: tuple = (0, 0, 0)
    wavelength: float = 16.0
    averaging_periods: float = 4.0
    map_axis: tuple = (0.0, 0.0, 1.0)
    name: str = "far field"

    def to_c(self) -> FarFieldSetupC:
        return FarFieldSetupC((C.c_int * 3)(*self.box_min), (C.c_int * 3)(*self.box_max), self.wavelength,
                              self.averaging_periods, (C.c_float * 3)(*self.map_axis))


@dataclass
class FluxSurface:
    """FluxSurface (src/FieldAnalyzer.hpp:32-51)."""
    kind: str = "plane"
    axis: int = 0
    index: int = 0
    rect_min: tuple = (0, 0)
    rect_max: tuple = (0, 0)
    box_min: tuple = (0, 0, 0)
    box_max: tuple = (0, 0, 0)
    label: str = "flux"

    @staticmethod
    def plane(axis, index, rect_min, rect_max, label="flux"):
        return FluxSurface("plane", max(0, min(2, axis)), index, tuple(rect_min), tuple(rect_max), label=label)

    @staticmethod
    def box(box_min, box_max, label="flux"):
        lo = tuple(min(a, b) for a, b in zip(box_min, box_max))
        hi = tuple(max(a, b) for a, b in zip(box_min, box_max))
        return FluxSurface("box", box_min=lo, box_max=hi, label=label)

    def to_c(self) -> FluxSurfaceC:
        return FluxSurfaceC(0 if self.kind == "plane" else 1, self.axis, self.index, (C.c_int * 2)(*self.rect_min),
                            (C.c_int * 2)(*self.rect_max), (C.c_int * 3)(*self.box_min), (C.c_int * 3)(*self.box_max))


@dataclass
class SubcellModel:
    """SubcellModel (src/SimulationGrid.hpp:96-101): edge overrides and thin-wire terms."""
    edges: list = field(default_factory=list)      # (texel(3), axis, sigma)
    wire_terms: list = field(default_factory=list)  # (texel(3), component, term, factor)

    def edge_array(self) -> np.ndarray:
        a = np.zeros(len(self.edges), EDGE_DTYPE)
        for i, (t, ax, s) in enumerate(self.edges):
            a[i] = (t, ax, s)
        return a

    def wire_array(self) -> np.ndarray:
        a = np.zeros(len(self.wire_terms), WIRE_DTYPE)
        for i, (t, comp, term, f) in enumerate(self.wire_terms):
            a[i] = (t, comp, term, f)
        return a


# ---------------------------------------------------------------------------------------------
# SimulationGrid and monitors
# ---------------------------------------------------------------------------------------------
class SimulationGrid:
    """EM_Lab's SimulationGrid on one HIP device (src/SimulationGrid.hpp:155-216)."""

    kMaxSources = 256

    def __init__(self, device: int = 0):
        lib = load_library()
        h = C.c_void_p()
        _check(lib.emfd_create(device, C.byref(h)))
        self._h = h
        self._lib = lib
        self._config: Optional[GridConfig] = None
        self._sources: list = []
        self._variant = KERNEL_AUTO
        self._storage = STORAGE_F32

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.emfd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def field_view(self) -> dict:
        """The grid's device planes as they are (emfd_grid_field_view): pointers valid until the
        next step or set-up call."""
        v = FieldViewC()
        _check(self._lib.emfd_grid_field_view(self._h, C.byref(v)))
        return dict(size=tuple(v.size), storage=v.storage, e=[v.e[i] for i in range(3)], b=[v.b[i] for i in range(3)],
                    scale=tuple(v.scale), sigma=v.sigma, eps_r=v.eps_r, device=v.device, step_count=v.step_count)

    # -- SimulationGrid API --
    def configure(self, config: GridConfig) -> GridConfig:
        applied = GridConfigC()
        _check(self._lib.emfd_grid_configure(self._h, C.byref(config.to_c()), C.byref(applied)))
        self._config = GridConfig.from_c(applied)
        self._absorber = ABSORBER_GRADED
        return self._config

    def upload_materials(self, sigma: np.ndarray, eps_r: np.ndarray, subcell: Optional[SubcellModel] = None) -> None:
        s = np.ascontiguousarray(sigma, dtype=np.float32).ravel()
        e = np.ascontiguousarray(eps_r, dtype=np.float32).ravel()
        subcell = subcell or SubcellModel()
        ed = subcell.edge_array()
        wi = subcell.wire_array()
        _check(self._lib.emfd_grid_upload_materials(self._h, _ptr(s), _ptr(e), s.size, _ptr(ed) if ed.size else None,
                                                    ed.size, _ptr(wi) if wi.size else None, wi.size))

    def set_sources(self, sources: Sequence[SourceDesc]) -> int:
        arr = (SourceDescC * max(len(sources), 1))(*[s.to_c() for s in sources])
        n = C.c_int(0)
        _check(self._lib.emfd_grid_set_sources(self._h, C.cast(arr, C.c_void_p), len(sources), C.byref(n)))
        self._sources = list(sources)
        return n.value

    def reset_fields(self) -> None:
        _check(self._lib.emfd_grid_reset_fields(self._h))

    def step(self, count: int) -> None:
        _check(self._lib.emfd_grid_step(self._h, int(count)))

    def synchronize(self) -> None:
        _check(self._lib.emfd_grid_synchronize(self._h))

    def set_kernel(self, variant: int) -> None:
        _check(self._lib.emfd_grid_set_kernel(self._h, int(variant)))
        self._variant = int(variant)

    def set_absorber(self, kind: int, sigma_max: float = 0.0, grading: float = 0.0, alpha_max: float = 0.0) -> None:
        """ABSORBER_GRADED (the reference's) or ABSORBER_CPML (convolutional PML in the same slabs;
        0 picks a parameter's default). Clears the fields and rewinds time."""
        _check(self._lib.emfd_grid_set_absorber(self._h, int(kind), float(sigma_max), float(grading),
                                                float(alpha_max)))
        self._absorber = int(kind)

    def set_storage(self, storage: int, scale_log2=None) -> None:
        """Field storage STORAGE_F32 / STORAGE_F16 / STORAGE_BF16 with a power-of-two scale per
        component (E x y z, B x y z). Clears the fields and rewinds time."""
        k = None if scale_log2 is None else (C.c_int * 6)(*[int(v) for v in scale_log2])
        _check(self._lib.emfd_grid_set_storage(self._h, int(storage), None if k is None else C.cast(k, C.c_void_p)))
        self._storage = int(storage)

    def fields_f64(self):
        """(E, B) as (3, n) float64 planes (exact for STORAGE_F64)."""
        nx, ny, nz = self.size
        n = nx * ny * nz
        E = np.empty((3, n), np.float64)
        B = np.empty((3, n), np.float64)
        _check(self._lib.emfd_grid_download_fields_f64(self._h, _ptr(E), _ptr(B)))
        return E, B

    def material_classes(self) -> int:
        c = C.c_int(0)
        _check(self._lib.emfd_grid_material_classes(self._h, C.byref(c)))
        return c.value

    def kernel_timing(self, enable: int = -1):
        """Live Yee-launch timing: enable = k > 0 brackets the Yee launch of every later step whose
        step count is a multiple of k with HIP events, 0 stops, -1 only reads. Returns the summed
        milliseconds and the number of bracketed launches since timing was switched on."""
        ms = C.c_double(0)
        n = C.c_int64(0)
        _check(self._lib.emfd_grid_kernel_timing(self._h, enable, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def kernel_timing_samples(self):
        """The bracketed Yee launches one by one (milliseconds, step order)."""
        n = C.c_int(0)
        _check(self._lib.emfd_grid_kernel_timing_samples(self._h, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), np.float32)
        _check(self._lib.emfd_grid_kernel_timing_samples(self._h, _ptr(out), out.size, C.byref(n)))
        return [float(v) for v in out[: n.value]]

    def yee_kernel_name(self) -> str:
        """The Yee kernel(s) a step launches with the current variant and storage."""
        v = self._variant & 0xFF
        if self._storage == STORAGE_F64:
            return "k_update_b_f64+k_update_e_f64"
        nx, ny, nz = self.size
        half = self._storage in (STORAGE_F16, STORAGE_BF16)
        cpml = getattr(self, "_absorber", ABSORBER_GRADED) == ABSORBER_CPML
        pml = self._config.pml_thickness if self._config is not None else 0
        if v == KERNEL_MARCH or (v == KERNEL_AUTO and not half and nx * ny * nz >= 8_000_000 and nx % 4 == 0
                                 and nx * ny * nz * 4 <= 0x80000000
                                 and (not cpml or min(nx, ny, nz) > 2 * pml + 2)):
            if cpml:
                return "k_yee_march (+k_march_fix_b, k_cpml_march_b, k_march_fix_e, k_cpml_march_e)"
            return "k_yee_march (+k_march_fix_b, k_march_fix_e)"
        if half:
            return "k_update_bh+k_update_eh"
        if v in (KERNEL_AUTO, KERNEL_TWO_PASS4) and self.size[0] % 4 == 0:
            return "k_update_b4+k_update_e4"
        return "k_update_b+k_update_e"

    @property
    def config(self) -> GridConfig:
        c = GridConfigC()
        _check(self._lib.emfd_grid_get_config(self._h, C.byref(c)))
        return GridConfig.from_c(c)

    @property
    def size(self) -> tuple:
        return tuple(self.config.size)

    def _state(self):
        t = C.c_double()
        n = C.c_uint64()
        dt = C.c_float()
        _check(self._lib.emfd_grid_state(self._h, C.byref(t), C.byref(n), C.byref(dt)))
        return t.value, n.value, dt.value

    @property
    def dt(self) -> float:
        return self._state()[2]

    @property
    def time(self) -> float:
        return self._state()[0]

    @property
    def step_count(self) -> int:
        return self._state()[1]

    def copy_probe(self, nbytes: int, iterations: int = 20) -> float:
        """Mean ms of a plain float4 device copy of nbytes (bandwidth / counter calibration)."""
        ms = C.c_double()
        _check(self._lib.emfd_copy_probe(self._h, int(nbytes), int(iterations), C.byref(ms)))
        return ms.value

    def stream_probe(self, plane_bytes: int, planes: int = 6, iterations: int = 5):
        """(mean ms, bytes moved) of the stream probe: planes in, planes out at one index."""
        ms = C.c_double()
        _check(self._lib.emfd_stream_probe(self._h, int(plane_bytes), int(planes), int(iterations), C.byref(ms)))
        unit = 256 * 2 * 4096 * 16
        moved = 2 * planes * (int(plane_bytes) // unit * unit)
        return ms.value, moved

    def gpu_memory_bytes(self) -> int:
        return int(self._lib.emfd_grid_memory_bytes(self._h))

    def fields(self, layout: int = LAYOUT_TEXEL):
        """(E, B) as numpy arrays: (n, 4) texels (alpha = sigma / eps_r) or (3, n) planes."""
        nx, ny, nz = self.size
        n = nx * ny * nz
        shape = (n, 4) if layout == LAYOUT_TEXEL else (3, n)
        E = np.empty(shape, np.float32)
        B = np.empty(shape, np.float32)
        _check(self._lib.emfd_grid_download_fields(self._h, _ptr(E), _ptr(B), layout))
        return E, B

    def set_fields(self, E: Optional[np.ndarray], B: Optional[np.ndarray], layout: int = LAYOUT_TEXEL) -> None:
        Ec = None if E is None else np.ascontiguousarray(E, np.float32)
        Bc = None if B is None else np.ascontiguousarray(B, np.float32)
        _check(self._lib.emfd_grid_upload_fields(self._h, _ptr(Ec), _ptr(Bc), layout))


@dataclass
class PortReading:
    valid: bool
    V: complex
    I: complex
    z_ohm: complex
    gamma: complex
    vswr: float
    return_loss_db: float
    accepted_power: float
    available_power: float


class LineMonitor:
    """The device side of CoaxMonitor (src/Coax.cpp:95-198): weighted sums of field texel
    components (quantities of terms), accumulated after every step into running-DFT phasors."""

    def __init__(self, grid: SimulationGrid):
        self.grid = grid
        self._lib = grid._lib
        self.n = 0

    def configure(self, quantities, terms, weights, wavelength: float, averaging_periods: float = 4.0) -> None:
        """quantities: (n, 4) int (first term, term count, 0 = E / 1 = B integral, unused);
        terms: (m, 4) int (x, y, z, component 0-2 E / 3-5 B); weights: (m,) float."""
        q = np.ascontiguousarray(np.asarray(quantities, np.int32).reshape(-1, 4))
        t = np.ascontiguousarray(np.asarray(terms, np.int32).reshape(-1, 4))
        w = np.ascontiguousarray(np.asarray(weights, np.float32).reshape(-1))
        if w.shape[0] != t.shape[0]:
            raise ValueError("one weight per term")
        _check(self._lib.emfd_lines_configure(self.grid.handle, _ptr(q) if q.size else None, q.shape[0],
                                              _ptr(t) if t.size else None, _ptr(w) if w.size else None, t.shape[0],
                                              float(wavelength), float(averaging_periods)))
        self.n = q.shape[0]

    def read(self):
        """(phasors (n, 4): stage-1 re, im, stage-2 re, im; instantaneous values (n,))."""
        ph = np.zeros((max(self.n, 1), 4), np.float32)
        now = np.zeros(max(self.n, 1), np.float32)
        _check(self._lib.emfd_lines_read(self.grid.handle, _ptr(ph), _ptr(now), ph.shape[0]))
        return ph[: self.n], now[: self.n]


def radar_backproject(grid: "SimulationGrid", setup: dict, z: np.ndarray, element_pos, weights, tx_pos,
                      tx_amp) -> np.ndarray:
    """Back-projection image (emfd_radar_backproject) on the grid's device: power [nr][nel][naz]
    (-1 outside the interior). setup keys: tx, axes (3 x 3 rows), interior_min, interior_max,
    wavelength, dt, sample_dt, r0, dr, nr, az0, daz, naz, el0, del, nel. z: [m][n_el] complex."""
    zz = np.ascontiguousarray(np.asarray(z, np.complex64)).view(np.float32)
    m, n_el = np.asarray(z).shape
    pos = np.ascontiguousarray(element_pos, np.float32).reshape(n_el, 3)
    w = np.ascontiguousarray(weights, np.float64).reshape(n_el)
    txp = np.ascontiguousarray(tx_pos, np.float32).reshape(-1, 3)
    amp = np.ascontiguousarray(tx_amp, np.float64).reshape(-1)
    f3 = C.c_float * 3
    st = RadarImageSetupC(f3(*setup["tx"]), (C.c_float * 9)(*np.asarray(setup["axes"], np.float32).ravel()),
                          f3(*setup["interior_min"]), f3(*setup["interior_max"]), setup["wavelength"], setup["dt"],
                          setup["sample_dt"], setup["r0"], setup["dr"], int(setup["nr"]), setup["az0"], setup["daz"],
                          int(setup["naz"]), setup["el0"], setup["del"], int(setup["nel"]))
    out = np.zeros((int(setup["nr"]), int(setup["nel"]), int(setup["naz"])), np.float64)
    _check(grid._lib.emfd_radar_backproject(grid.handle, C.byref(st), _ptr(zz), int(m), int(n_el), _ptr(pos), _ptr(w),
                                            _ptr(txp) if txp.size else None, _ptr(amp) if amp.size else None,
                                            int(amp.size), _ptr(out)))
    return out


def coax_geometry_factor(inner: int, outer: int) -> float:
    """CoaxMonitor::geometryFactor (src/Coax.cpp:42-83): lattice capacitance per unit length of the
    square coax in units of eps; Z0 = eta0 / (sqrt(eps_r) G)."""
    return float(load_library().emfd_coax_geometry_factor(int(inner), int(outer)))


class PortMonitor:
    """PortMonitor on a grid's context (src/Ports.hpp:84-140)."""

    def __init__(self, grid: SimulationGrid):
        self.grid = grid
        self._lib = grid._lib
        self.ports: list = []
        self._sweep: Optional[SweepSetup] = None

    def configure(self, ports: Sequence[PortSetup], averaging_periods: float, sweep: Optional[SweepSetup] = None,
                  cell_size: float = 1.0e-3, record_steps: int = 0) -> None:
        arr = (PortSetupC * max(len(ports), 1))(*[p.to_c() for p in ports])
        sw = SweepSetupC(sweep.wavelength_min, sweep.wavelength_max, sweep.points) if sweep else None
        _check(self._lib.emfd_ports_configure(self.grid.handle, C.cast(arr, C.c_void_p), len(ports),
                                              averaging_periods, C.byref(sw) if sw else None, cell_size,
                                              int(record_steps)))
        self.ports = list(ports)[:256]
        self._sweep = sweep

    def disable(self) -> None:
        _check(self._lib.emfd_ports_disable(self.grid.handle))
        self.ports = []

    def phasors(self) -> np.ndarray:
        out = np.zeros((max(len(self.ports), 1), 8), np.float32)
        _check(self._lib.emfd_ports_read_phasors(self.grid.handle, _ptr(out), out.shape[0]))
        return out[: len(self.ports)]

    def set_dft(self, wavelengths) -> None:
        """FP64 running DFTs of every port's V and I at these wavelengths (empty: off)."""
        w = np.ascontiguousarray(wavelengths, np.float64).ravel()
        _check(self._lib.emfd_ports_set_dft(self.grid.handle, _ptr(w) if w.size else None, w.size))
        self.dft_wavelengths = [float(v) for v in w]

    def dft(self) -> np.ndarray:
        """(n_ports, n_wavelengths, 2) complex128: V(w), I(w)."""
        nw = len(getattr(self, "dft_wavelengths", []))
        out = np.zeros((len(self.ports), nw, 4), np.float64)
        _check(self._lib.emfd_ports_read_dft(self.grid.handle, _ptr(out), out.size))
        return out[..., 0::2] + 1j * out[..., 1::2]

    def evaluate(self) -> list:
        n = len(self.ports)
        arr = (PortReadingC * max(n, 1))()
        _check(self._lib.emfd_ports_evaluate(self.grid.handle, C.cast(arr, C.c_void_p), n))
        res = []
        for r in arr[:n]:
            res.append(PortReading(bool(r.valid), complex(r.v_re, r.v_im), complex(r.i_re, r.i_im),
                                   complex(r.z_re, r.z_im), complex(r.gamma_re, r.gamma_im), r.vswr,
                                   r.return_loss_db, r.accepted_power, r.available_power))
        return res

    def recorded_steps(self) -> int:
        return int(self._lib.emfd_ports_recorded_steps(self.grid.handle))

    def read_series(self) -> np.ndarray:
        n = self.recorded_steps()
        out = np.zeros((n, len(self.ports), 2), np.float32)
        _check(self._lib.emfd_ports_read_series(self.grid.handle, _ptr(out), out.size))
        return out

    def evaluate_sweep(self, dt: float) -> dict:
        n = len(self.ports)
        pts = max(self._sweep.points if self._sweep else 2, 2)
        parr = (SweepPointC * (n * pts))()
        sarr = (PortSweepC * max(n, 1))()
        res = C.c_double()
        samples = C.c_int64()
        full = C.c_int()
        _check(self._lib.emfd_ports_evaluate_sweep(self.grid.handle, dt, C.cast(parr, C.c_void_p),
                                                   C.cast(sarr, C.c_void_p), C.byref(res), C.byref(samples),
                                                   C.byref(full)))
        ports = []
        for p in range(n):
            points = np.array([(q.wavelength, q.z_re, q.z_im, q.s11_db, q.vswr) for q in parr[p * pts:(p + 1) * pts]])
            s = sarr[p]
            ports.append(dict(points=points, best=s.best, best_wavelength=s.best_wavelength, has_band=bool(s.has_band),
                              band_low_wavelength=s.band_low_wavelength, band_high_wavelength=s.band_high_wavelength))
        return dict(residual=res.value, samples=samples.value, full=bool(full.value), ports=ports)


class FarFieldMonitor:
    """FarFieldMonitor on a grid's context (src/FarField.hpp:75-133)."""

    def __init__(self, grid: SimulationGrid):
        self.grid = grid
        self._lib = grid._lib
        self.setup: Optional[FarFieldSetup] = None
        self.n_points = 0

    def configure(self, setup: FarFieldSetup) -> FarFieldSetup:
        applied = FarFieldSetupC()
        npts = C.c_int()
        _check(self._lib.emfd_ntff_configure(self.grid.handle, C.byref(setup.to_c()), C.byref(applied),
                                             C.byref(npts)))
        self.setup = FarFieldSetup(tuple(applied.box_min), tuple(applied.box_max), applied.wavelength,
                                   applied.averaging_periods, tuple(applied.map_axis), setup.name)
        self.n_points = npts.value
        return self.setup

    def disable(self) -> None:
        _check(self._lib.emfd_ntff_disable(self.grid.handle))

    def reset_accumulation(self) -> None:
        _check(self._lib.emfd_ntff_reset(self.grid.handle))

    def phasors(self, freq: int = 0) -> np.ndarray:
        out = np.zeros((self.n_points, 32), np.float32)
        _check(self._lib.emfd_ntff_read_phasors_at(self.grid.handle, int(freq), _ptr(out), self.n_points))
        return out

    def set_wavelengths(self, wavelengths) -> None:
        """Accumulate at several wavelengths (the first becomes the monitor's wavelength)."""
        w = np.ascontiguousarray(wavelengths, np.float32)
        _check(self._lib.emfd_ntff_set_wavelengths(self.grid.handle, _ptr(w), w.size))
        self.wavelengths = [float(max(v, 2.0)) for v in w]
        if self.setup is not None:
            self.setup.wavelength = self.wavelengths[0]

    def directions(self):
        n = C.c_int()
        _check(self._lib.emfd_ntff_directions(self.grid.handle, None, 0, C.byref(n), None, None))
        dirs = np.zeros((n.value, 4), np.float32)
        mi = np.zeros(2 * 81 * 81, np.int32)
        frame = np.zeros(9, np.float32)
        _check(self._lib.emfd_ntff_directions(self.grid.handle, _ptr(dirs), n.value, C.byref(n), _ptr(mi),
                                              _ptr(frame)))
        return dirs, mi, frame

    def radiation(self, dirs: np.ndarray, freq: int = 0):
        """|N_perp - r x L|^2 per direction and the radiated power P (wavelength `freq` of the list)."""
        d = np.ascontiguousarray(dirs, np.float32)
        if d.ndim == 2 and d.shape[1] == 3:
            d = np.ascontiguousarray(np.concatenate([d, np.zeros((d.shape[0], 1), np.float32)], 1))
        out = np.zeros(d.shape[0], np.float32)
        P = C.c_double()
        _check(self._lib.emfd_ntff_eval_at(self.grid.handle, int(freq), _ptr(d), d.shape[0], _ptr(out), C.byref(P)))
        return out, P.value

    def set_dft(self, enable: bool = True) -> None:
        """FP64 running DFTs of the Huygens-node fields at every far-field wavelength."""
        _check(self._lib.emfd_ntff_set_dft(self.grid.handle, 1 if enable else 0))

    def dft(self, freq: int = 0):
        """(E, H) running DFTs of wavelength `freq`: two (n_points, 3) complex128 arrays."""
        out = np.zeros((self.n_points, 12), np.float64)
        _check(self._lib.emfd_ntff_read_dft(self.grid.handle, int(freq), _ptr(out), self.n_points))
        return out[:, 0:3] + 1j * out[:, 3:6], out[:, 6:9] + 1j * out[:, 9:12]

    def radiation_dft(self, dirs: np.ndarray, freq: int = 0):
        """radiation() on the running DFTs."""
        d = np.ascontiguousarray(dirs, np.float32)
        if d.ndim == 2 and d.shape[1] == 3:
            d = np.ascontiguousarray(np.concatenate([d, np.zeros((d.shape[0], 1), np.float32)], 1))
        out = np.zeros(d.shape[0], np.float32)
        P = C.c_double()
        _check(self._lib.emfd_ntff_eval_dft(self.grid.handle, int(freq), _ptr(d), d.shape[0], _ptr(out), C.byref(P)))
        return out, P.value

    @staticmethod
    def theta_phi_grid(n_theta: int, n_phi: int) -> np.ndarray:
        """Directions of a (theta, phi) grid, theta in [0, 180], phi in [0, 360] degrees inclusive,
        theta-major: (n_theta * n_phi, 4) float32 (x, y, z, 0)."""
        th = np.radians(np.linspace(0.0, 180.0, n_theta))
        ph = np.radians(np.linspace(0.0, 360.0, n_phi))
        t, p = np.meshgrid(th, ph, indexing="ij")
        d = np.stack([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t), np.zeros_like(t)], -1)
        return np.ascontiguousarray(d.reshape(-1, 4), np.float32)

    def evaluate(self) -> dict:
        dirs, _, _ = self.directions()
        valid = C.c_int()
        P = C.c_double()
        mx = C.c_float()
        mdir = np.zeros(3, np.float32)
        dbi = np.zeros(dirs.shape[0], np.float32)
        cuts = np.zeros(18, np.float32)
        _check(self._lib.emfd_ntff_evaluate(self.grid.handle, C.byref(valid), C.byref(P), C.byref(mx), _ptr(mdir),
                                            _ptr(dbi), _ptr(cuts)))
        names = ("xy", "xz", "yz")
        return dict(valid=bool(valid.value), radiated_power=P.value, max_directivity_dbi=mx.value,
                    max_direction=mdir, directivity_dbi=dbi,
                    cuts={names[c]: dict(peak_dbi=float(cuts[6 * c]), peak_angle_deg=float(cuts[6 * c + 1]),
                                         beamwidth_deg=float(cuts[6 * c + 2]), front_to_back_db=float(cuts[6 * c + 3]))
                          for c in range(3)})


# ---------------------------------------------------------------------------------------------
# Report files (PortMonitor::writeSweepCsv, src/Ports.cpp:280-297; FarFieldMonitor::writeCsv,
# src/FarField.cpp:337-350): the same columns, numbers printed as C++ streams print them (%g).
# ---------------------------------------------------------------------------------------------
C0 = 299792458.0  # src/Ports.cpp kC0


def _g(v) -> str:
    return format(float(v), "g")


def write_sweep_csv(path, port_names, sweep: dict, cell_size: float) -> bool:
    """Sweep CSV: frequency_Hz, wavelength_cells, then R, X, S11 dB, VSWR per port; `sweep` is
    PortMonitor.evaluate_sweep()'s (or Batch.sweep_result()'s) dictionary."""
    ports = sweep.get("ports") or []
    if not ports:
        return False
    lines = ["frequency_Hz,wavelength_cells" +
             "".join(f",{n}_R_ohm,{n}_X_ohm,{n}_S11_dB,{n}_VSWR" for n in port_names)]
    pts0 = ports[0]["points"]
    for i in range(len(pts0)):
        wl = float(pts0[i][0])
        row = [_g(C0 / (wl * cell_size)), _g(wl)]
        for p in ports:
            q = p["points"][i]
            row += [_g(q[1]), _g(q[2]), _g(q[3]), _g(q[4])]
        lines.append(",".join(row))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return True


def write_pattern_csv(path, pattern: dict) -> bool:
    """Far-field cut CSV: angle_deg, then xy_dBi, xz_dBi, yz_dBi; `pattern` is
    FarFieldMonitor.evaluate()'s dictionary (cut c holds directions 180 c .. 180 c + 179)."""
    if not pattern.get("valid"):
        return False
    d = np.asarray(pattern["directivity_dbi"], np.float32)
    names = ("xy", "xz", "yz")
    lines = ["angle_deg" + "".join(f",{n}_dBi" for n in names)]
    for i in range(180):
        angle = np.float32(360.0) * np.float32(i) / np.float32(180)
        lines.append(",".join([_g(angle)] + [_g(d[180 * c + i]) for c in range(3)]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return True


class FieldAnalyzer:
    """FieldAnalyzer probes and flux meters on a grid's context (src/FieldAnalyzer.hpp:73-130)."""

    def __init__(self, grid: SimulationGrid):
        self.grid = grid
        self._lib = grid._lib

    def measure_flux(self, surfaces) -> np.ndarray:
        if isinstance(surfaces, FluxSurface):
            surfaces = [surfaces]
        arr = (FluxSurfaceC * max(len(surfaces), 1))(*[s.to_c() for s in surfaces])
        out = np.zeros(max(len(surfaces), 1), np.float64)
        _check(self._lib.emfd_flux_measure(self.grid.handle, C.cast(arr, C.c_void_p), len(surfaces), _ptr(out)))
        return out[: len(surfaces)]

    def probe(self, positions) -> list:
        p = np.ascontiguousarray(np.atleast_2d(np.asarray(positions, np.float32)))
        arr = (ProbeSampleC * max(p.shape[0], 1))()
        _check(self._lib.emfd_probe_measure(self.grid.handle, _ptr(p), p.shape[0], C.cast(arr, C.c_void_p)))
        return [dict(position=np.array(s.position), E=np.array(s.E), B=np.array(s.B), S=np.array(s.S),
                     sigma=s.sigma, epsilon_r=s.epsilon_r) for s in arr[: p.shape[0]]]


class Camera:
    """EM_Lab's orbit camera (src/Camera.cpp:29-36,54-68,102-110, defaults src/Camera.hpp:48-58):
    eye = target + distance (cos p cos y, cos p sin y, sin p), looking at the target, z up;
    glm's right-handed lookAt and perspective (clip z in -1 .. 1)."""

    def __init__(self, target=(0.0, 0.0, 0.0), distance=4.0, yaw_deg=0.0, pitch_deg=0.0, fov_deg=50.0,
                 aspect=16.0 / 9.0, near=0.01, far=100.0):
        self.target = np.asarray(target, np.float64)
        self.distance = float(np.clip(distance, 0.05, 50.0))
        self.yaw = math.radians(yaw_deg)
        self.pitch = float(np.clip(math.radians(pitch_deg), -math.radians(89.0), math.radians(89.0)))
        self.fov = math.radians(float(np.clip(fov_deg, 10.0, 120.0)))
        self.aspect = float(aspect)
        self.near = max(float(near), 1e-4)
        self.far = max(float(far), self.near * 2.0)

    def position(self) -> np.ndarray:
        cp = math.cos(self.pitch)
        d = np.array([cp * math.cos(self.yaw), cp * math.sin(self.yaw), math.sin(self.pitch)])
        return self.target + self.distance * d

    def view(self) -> np.ndarray:
        eye = self.position()
        f = self.target - eye
        f = f / np.linalg.norm(f)
        s = np.cross(f, np.array([0.0, 0.0, 1.0]))
        s = s / np.linalg.norm(s)
        u = np.cross(s, f)
        m = np.eye(4)
        m[0, :3], m[1, :3], m[2, :3] = s, u, -f
        m[0, 3], m[1, 3], m[2, 3] = -s @ eye, -u @ eye, f @ eye
        return m

    def projection(self) -> np.ndarray:
        t = math.tan(self.fov / 2.0)
        n, fa = self.near, self.far
        m = np.zeros((4, 4))
        m[0, 0] = 1.0 / (self.aspect * t)
        m[1, 1] = 1.0 / t
        m[2, 2] = -(fa + n) / (fa - n)
        m[2, 3] = -2.0 * fa * n / (fa - n)
        m[3, 2] = -1.0
        return m

    def inverse_view_projection(self) -> np.ndarray:
        """16 floats, column-major (glm's layout): element (r, c) at 4 c + r."""
        inv = np.linalg.inv(self.projection() @ self.view())
        return np.ascontiguousarray(inv.T.ravel(), dtype=np.float32)


RENDER_FIELDS, RENDER_E, RENDER_B, RENDER_POYNTING, RENDER_SIGNED_E = 0, 1, 2, 3, 4
IMAGE_RGBA8, IMAGE_RGBA32F, IMAGE_DEVICE = 0, 1, 0x100


class RenderSettings:
    """RenderSettings (src/VolumeRenderer.hpp:21-39) with the reference's defaults."""

    def __init__(self, **kw):
        c = RenderSettingsC()
        load_library().emfd_render_default_settings(C.byref(c))
        self.mode, self.signed_axis = c.mode, c.signed_axis
        self.gain_e, self.gain_b, self.density, self.emission = c.gain_e, c.gain_b, c.density, c.emission
        self.step_voxels, self.material_opacity, self.exposure = c.step_voxels, c.material_opacity, c.exposure
        self.clip_min, self.clip_max = tuple(c.clip_min), tuple(c.clip_max)
        for k, v in kw.items():
            if not hasattr(self, k):
                raise TypeError(f"unknown render setting {k}")
            setattr(self, k, v)

    def to_c(self) -> RenderSettingsC:
        c = RenderSettingsC()
        c.mode, c.signed_axis = int(self.mode), int(self.signed_axis)
        c.gain_e, c.gain_b, c.density, c.emission = self.gain_e, self.gain_b, self.density, self.emission
        c.step_voxels, c.material_opacity, c.exposure = self.step_voxels, self.material_opacity, self.exposure
        c.clip_min[:] = [float(v) for v in self.clip_min]
        c.clip_max[:] = [float(v) for v in self.clip_max]
        return c


class VolumeRenderer:
    """The viewer's field raymarch over a grid's device planes (VolumeRenderer::render,
    src/VolumeRenderer.cpp:70-121; shaders/raymarch.frag)."""

    def render(self, grid: "SimulationGrid", camera: Camera, settings: Optional[RenderSettings] = None,
               width: int = 320, height: int = 180, float_image: bool = False) -> np.ndarray:
        """(height, width, 4) image, bottom row first: uint8 RGBA, or the shader's float output."""
        settings = settings or RenderSettings()
        ivp = camera.inverse_view_projection()
        eye = np.ascontiguousarray(camera.position(), dtype=np.float32)
        out = np.zeros((height, width, 4), np.float32 if float_image else np.uint8)
        sc = settings.to_c()
        _check(grid._lib.emfd_render_volume(grid.handle, C.byref(sc), _ptr(ivp), _ptr(eye), int(width), int(height),
                                            IMAGE_RGBA32F if float_image else IMAGE_RGBA8, _ptr(out)))
        return out


class Batch:
    """Batched stepper: many independent scenes advance together (emfd_batch_* in emfdtd.h).

    Each scene is what SimulationGrid.configure + upload_materials + set_sources and
    PortMonitor.configure (with an optional sweep) would set up alone, and gives bit-identical
    fields, port phasors and series."""

    def __init__(self, device: int = 0):
        lib = load_library()
        h = C.c_void_p()
        _check(lib.emfd_batch_create(device, C.byref(h)))
        self._h = h
        self._lib = lib
        self.configs: list = []
        self.ports: list = []
        self.sweeps: list = []
        self._keep: list = []

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.emfd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, config: GridConfig, sigma: Optional[np.ndarray] = None, eps_r: Optional[np.ndarray] = None,
                  subcell: Optional[SubcellModel] = None, sources: Sequence[SourceDesc] = (),
                  ports: Sequence[PortSetup] = (), averaging_periods: float = 4.0,
                  sweep: Optional[SweepSetup] = None) -> int:
        sc = BatchSceneC()
        sc.config = config.to_c()
        keep = []
        if sigma is not None:
            sg = np.ascontiguousarray(sigma, np.float32).ravel()
            ep = np.ascontiguousarray(eps_r, np.float32).ravel()
            keep += [sg, ep]
            sc.sigma, sc.eps_r, sc.n_cells = _ptr(sg), _ptr(ep), sg.size
        edges = subcell.edge_array() if subcell else np.zeros(0, EDGE_DTYPE)
        wires = subcell.wire_array() if subcell else np.zeros(0, WIRE_DTYPE)
        keep += [edges, wires]
        sc.edges, sc.n_edges = (_ptr(edges) if edges.size else None), edges.size
        sc.wires, sc.n_wires = (_ptr(wires) if wires.size else None), wires.size
        src = (SourceDescC * max(len(sources), 1))(*[x.to_c() for x in sources])
        prt = (PortSetupC * max(len(ports), 1))(*[x.to_c() for x in ports])
        keep += [src, prt]
        sc.sources, sc.n_sources = C.cast(src, C.c_void_p), len(sources)
        sc.ports, sc.n_ports = C.cast(prt, C.c_void_p), len(ports)
        sc.averaging_periods = averaging_periods
        if sweep is not None:
            sw = SweepSetupC(sweep.wavelength_min, sweep.wavelength_max, sweep.points)
            keep.append(sw)
            sc.sweep = C.cast(C.pointer(sw), C.c_void_p)
        idx = C.c_int()
        applied = GridConfigC()
        _check(self._lib.emfd_batch_add_scene(self._h, C.byref(sc), C.byref(idx), C.byref(applied)))
        self.configs.append(GridConfig.from_c(applied))
        self.ports.append(list(ports))
        self.sweeps.append(sweep)
        return idx.value

    def add(self, scene) -> int:
        """Adds an em_lab_amd.scene.Scene (its thin wires must already be emitted)."""
        s, e = scene.flat_materials()
        vacuum = not np.any(s) and np.all(e == 1.0)
        return self.add_scene(scene.grid_config(), None if vacuum else s, None if vacuum else e, scene.subcell,
                              scene.sources, scene.ports,
                              scene.far_field.averaging_periods if scene.far_field else 4.0, scene.sweep)

    def finalize(self, max_steps: int) -> None:
        _check(self._lib.emfd_batch_finalize(self._h, int(max_steps)))

    def info(self) -> dict:
        n, cells, tiles, nbytes = C.c_int(), C.c_int64(), C.c_int64(), C.c_size_t()
        _check(self._lib.emfd_batch_info(self._h, C.byref(n), C.byref(cells), C.byref(tiles), C.byref(nbytes)))
        return dict(scenes=n.value, cells=cells.value, tiles=tiles.value, device_bytes=nbytes.value)

    def set_launch(self, quads_per_thread: int = 1, persistent_blocks: int = 0, xcd_slices: int = 8) -> None:
        _check(self._lib.emfd_batch_set_launch(self._h, C.byref(BatchLaunchC(quads_per_thread, persistent_blocks,
                                                                               xcd_slices))))

    def launch(self) -> tuple:
        l = BatchLaunchC()
        _check(self._lib.emfd_batch_get_launch(self._h, C.byref(l)))
        return (l.quads_per_thread, l.persistent_blocks, l.xcd_slices)

    def autotune(self, iterations: int = 5):
        chosen = BatchLaunchC()
        cap = 64
        table = np.zeros(3 * cap, np.float64)
        n = C.c_int()
        _check(self._lib.emfd_batch_autotune(self._h, iterations, C.byref(chosen), _ptr(table), cap, C.byref(n)))
        return (chosen.quads_per_thread, chosen.persistent_blocks, chosen.xcd_slices), table[:3 * n.value].reshape(-1, 3)

    def set_graph(self, steps_per_graph: int) -> None:
        _check(self._lib.emfd_batch_set_graph(self._h, int(steps_per_graph)))

    def step(self, count: int) -> None:
        _check(self._lib.emfd_batch_step(self._h, int(count)))

    def synchronize(self) -> None:
        _check(self._lib.emfd_batch_synchronize(self._h))

    @property
    def step_count(self) -> int:
        n = C.c_uint64()
        _check(self._lib.emfd_batch_state(self._h, C.byref(n)))
        return n.value

    def reset(self) -> None:
        _check(self._lib.emfd_batch_reset(self._h))

    def fields(self, scene: int):
        nx, ny, nz = self.configs[scene].size
        n = nx * ny * nz
        E = np.empty(3 * n, np.float32)
        B = np.empty(3 * n, np.float32)
        _check(self._lib.emfd_batch_download_fields(self._h, scene, _ptr(E), _ptr(B)))
        return E, B

    def set_fields(self, scene: int, E: Optional[np.ndarray], B: Optional[np.ndarray]) -> None:
        e = None if E is None else np.ascontiguousarray(E, np.float32).ravel()
        b = None if B is None else np.ascontiguousarray(B, np.float32).ravel()
        _check(self._lib.emfd_batch_upload_fields(self._h, scene, _ptr(e), _ptr(b)))

    def phasors(self, scene: int) -> np.ndarray:
        n = len(self.ports[scene])
        out = np.zeros(8 * max(n, 1), np.float32)
        _check(self._lib.emfd_batch_read_phasors(self._h, scene, _ptr(out), out.size))
        return out[:8 * n].reshape(n, 8)

    def series(self, scene: int) -> np.ndarray:
        n = len(self.ports[scene])
        out = np.zeros(2 * n * max(self.step_count, 1), np.float32)
        _check(self._lib.emfd_batch_read_series(self._h, scene, _ptr(out), out.size))
        return out[:2 * n * self.step_count].reshape(self.step_count, n, 2)

    def evaluate_sweeps(self) -> None:
        _check(self._lib.emfd_batch_evaluate_sweeps(self._h))

    def sweep_result(self, scene: int) -> dict:
        n = len(self.ports[scene])
        pts = max(self.sweeps[scene].points, 2)
        parr = (SweepPointC * (n * pts))()
        sarr = (PortSweepC * max(n, 1))()
        res = C.c_double()
        _check(self._lib.emfd_batch_sweep_result(self._h, scene, C.cast(parr, C.c_void_p), n * pts,
                                                  C.cast(sarr, C.c_void_p), C.byref(res)))
        ports = []
        for p in range(n):
            points = np.array([(q.wavelength, q.z_re, q.z_im, q.s11_db, q.vswr) for q in parr[p * pts:(p + 1) * pts]])
            s = sarr[p]
            ports.append(dict(points=points, best=s.best, best_wavelength=s.best_wavelength, has_band=bool(s.has_band),
                              band_low_wavelength=s.band_low_wavelength, band_high_wavelength=s.band_high_wavelength))
        return dict(residual=res.value, ports=ports)

    def kernel_timing(self, enable: int = -1):
        ms = C.c_double()
        n = C.c_int64()
        _check(self._lib.emfd_batch_kernel_timing(self._h, enable, C.byref(ms), C.byref(n)))
        return ms.value, n.value


class Mlp:
    """The surrogate network on one HIP device (emfd_mlp_* in emfdtd.h): BF16 MFMA forward and
    backward, FP32 master parameters, Adam. Parameters: per layer W [n_out][n_in], then b [n_out]."""

    def __init__(self, dims: Sequence[int], max_batch: int = 65536, device: int = 0):
        lib = load_library()
        h = C.c_void_p()
        d = (C.c_int * len(dims))(*[int(v) for v in dims])
        _check(lib.emfd_mlp_create(device, C.cast(d, C.c_void_p), len(dims), int(max_batch), C.byref(h)))
        self._h = h
        self._lib = lib
        self.dims = tuple(int(v) for v in dims)
        self.max_batch = int(max_batch)
        n = C.c_int64()
        _check(lib.emfd_mlp_param_count(h, C.byref(n)))
        self.n_params = n.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.emfd_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_params(self, seed: int = 0) -> np.ndarray:
        """He-style uniform initialisation (seeded), biases zero; sets and returns the vector."""
        rng = np.random.default_rng(seed)
        parts = []
        for i, o in zip(self.dims[:-1], self.dims[1:]):
            lim = math.sqrt(6.0 / i)
            parts.append(rng.uniform(-lim, lim, size=(o, i)).astype(np.float32).ravel())
            parts.append(np.zeros(o, np.float32))
        p = np.concatenate(parts)
        self.set_params(p)
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        _check(self._lib.emfd_mlp_set_params(self._h, _ptr(p), p.size))

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        _check(self._lib.emfd_mlp_get_params(self._h, _ptr(p), p.size))
        return p

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.dims[-1]), np.float32)
        _check(self._lib.emfd_mlp_forward(self._h, _ptr(x), x.shape[0], _ptr(y)))
        return y

    def gradients(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_double()
        _check(self._lib.emfd_mlp_gradients(self._h, _ptr(x), _ptr(t), x.shape[0], _ptr(g), C.byref(loss)))
        return loss.value, g

    def set_adam(self, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0) -> None:
        _check(self._lib.emfd_mlp_set_adam(self._h, beta1, beta2, eps, weight_decay))

    def train_step(self, x: np.ndarray, t: np.ndarray, lr: float) -> float:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_double()
        _check(self._lib.emfd_mlp_train_step(self._h, _ptr(x), _ptr(t), x.shape[0], lr, C.byref(loss)))
        return loss.value

    def set_dataset(self, x: np.ndarray, t: np.ndarray) -> None:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        _check(self._lib.emfd_mlp_set_dataset(self._h, _ptr(x), _ptr(t), x.shape[0]))

    def train(self, batch: int, epochs: int, lr_start: float, lr_end: float, seed: int = 0) -> np.ndarray:
        losses = np.zeros(max(epochs, 1), np.float64)
        _check(self._lib.emfd_mlp_train(self._h, batch, epochs, lr_start, lr_end, seed, _ptr(losses), losses.size))
        return losses[:epochs]

    def time_forward(self, batch: int, iterations: int = 20) -> float:
        ms = C.c_double()
        _check(self._lib.emfd_mlp_time_forward(self._h, batch, iterations, C.byref(ms)))
        return ms.value


def version() -> str:
    return load_library().emfd_version().decode()


def dipole_source(center, axis, wavelength, current, length=0.0) -> SourceDesc:
    """SourceDesc::dipole (src/SimulationGrid.cpp:70-82), float arithmetic as the reference."""
    f = np.float32
    soft = f(0.75)
    amp = f(current) / (f(math.pi) * soft * soft)
    ax = np.asarray(axis, np.float32)
    len2 = f(np.dot(ax, ax))
    n = ax / np.sqrt(len2) if len2 > 0 else np.array([0, 0, 1], np.float32)
    half = np.abs(n) * (f(0.5) * f(max(length, 0.0)))
    return SourceDesc(center=tuple(float(v) for v in center), half_extent=tuple(float(v) for v in half),
                      softness=(0.75, 0.75, 0.75), polarization=tuple(float(v) for v in ax), amplitude=float(amp),
                      wavelength=float(wavelength))
