"""Build libemfdtd.so (HIP kernels + C ABI) for gfx950, in-tree.

Run as ``python -m em_lab_amd.build``. The library lands in ``em_lab_amd/_lib/libemfdtd.so``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB_DIR = PKG / "_lib"
LIB = LIB_DIR / "libemfdtd.so"
OBJ_DIR = PKG / "_build"

SOURCES = ["kernels.hip", "yee_two_pass4.hip", "yee_march.hip", "yee_half.hip", "yee_f64.hip", "monitors.hip", "farfield.hip", "dft.hip", "noalloc.hip", "cpml.hip", "batch.hip", "mlp.hip", "radar.hip", "render.hip",
           "grid.cpp", "monitors.cpp", "batch.cpp", "mlp.cpp", "radar.cpp", "render.cpp"]

# -ffp-contract=off: every expression rounds as the reference shader writes it (no fused
# multiply-add), so results can be compared against the CPU restatement to the last ulp
# outside exp/sin/cos. Packed FP32 instructions are disabled (-packed-fp32-ops).
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libemfdtd.so")


def build(verbose: bool = False) -> Path:
    """Compile every source and link the shared library. Always rebuilds."""
    cc = hipcc()
    OBJ_DIR.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = OBJ_DIR / (src.replace(".", "_") + ".o")
        cmd = [cc, *HIP_FLAGS]
        if src.endswith(".cpp"):
            cmd += ["-x", "hip"]
        cmd += ["-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append(f"--- {src}\n{out}")
        elif verbose and out.strip():
            print(out)
    if failed:
        raise RuntimeError("libemfdtd build failed:\n" + "\n".join(failed))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("libemfdtd link failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    root = PKG.parent
    r = subprocess.run(["make", "-B", "-s", "-C", str(root / "oracle")], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    return root / "oracle" / "_build" / "liboracle.so"


INTEGRATION_SRC = PKG / "integration" / "simulation_grid.cpp"
INTEGRATION_BIN = LIB_DIR / "simulation_grid"


def build_integration() -> Path:
    """Compile the C++ binding example (a host program calling the C ABI) against
    include/emfdtd.h and libemfdtd.so; it finds the library next to itself."""
    cxx = shutil.which("g++") or shutil.which("c++")
    if not cxx:
        raise RuntimeError("no host C++ compiler for the binding example")
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", str(PKG.parent / "include"), str(INTEGRATION_SRC),
           "-o", str(INTEGRATION_BIN), "-L", str(LIB_DIR), "-lemfdtd", "-Wl,-rpath,$ORIGIN",
           "-Wl,--allow-shlib-undefined"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("binding example build failed:\n" + r.stdout)
    return INTEGRATION_BIN


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_oracle())
    print(build_integration())
