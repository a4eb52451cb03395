"""CPU-side checks of the C ABI library: it loads and exports every entry point that
include/emfdtd.h declares (no compute calls: there is no GPU here)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "emfdtd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(emfd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_path():
    names = declared_functions()
    for must in ["emfd_create", "emfd_grid_configure", "emfd_grid_upload_materials", "emfd_grid_set_sources",
                 "emfd_grid_step", "emfd_ports_configure", "emfd_ports_evaluate_sweep", "emfd_ntff_configure",
                 "emfd_ntff_evaluate", "emfd_flux_measure", "emfd_probe_measure"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib_path = ROOT / "em_lab_amd" / "_lib" / "libemfdtd.so"
    if not lib_path.exists():
        from em_lab_amd import build
        build.build()
    lib = ctypes.CDLL(str(lib_path))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    lib.emfd_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.emfd_version()


def test_python_binding_signatures_cover_header():
    from em_lab_amd import emfdtd
    assert set(declared_functions()) == set(emfdtd._SIGNATURES)


def test_binding_example_builds_and_links():
    """The integration guide's C++ binding compiles against the header, links the library and
    runs up to its argument check (no device call)."""
    import subprocess
    from em_lab_amd import build
    lib_path = ROOT / "em_lab_amd" / "_lib" / "libemfdtd.so"
    if not lib_path.exists():
        build.build()
    exe = build.build_integration()
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
