"""em_lab_amd — MI355X-native FDTD engine behind EM_Lab's solver interfaces.

The compute path is libemfdtd.so (hand-written HIP kernels for gfx950 behind a C ABI,
include/emfdtd.h); this package holds its Python host binding and the scene builder.
"""
from .emfdtd import (FarFieldMonitor, FarFieldSetup, FieldAnalyzer, FluxSurface, GridConfig, PortMonitor,
                     PortSetup, SimulationGrid, SourceDesc, SubcellModel, SweepSetup, load_library, version)

__all__ = ["FarFieldMonitor", "FarFieldSetup", "FieldAnalyzer", "FluxSurface", "GridConfig", "PortMonitor",
           "PortSetup", "SimulationGrid", "SourceDesc", "SubcellModel", "SweepSetup", "load_library", "version"]
