"""The oracle's pins reject plausible mistakes.

Each case below plants one mistake in a copy of the oracle's sources (oracle/oracle.c,
oracle/render.c: a transposed neighbour, a wrong component, a half-cell shift, a sign), compiles
the copies into a temporary library with the same flags as oracle/Makefile, and runs the pins of
tests/test_oracle_pins.py and tests/test_render_cpu.py that cover the mutated function against it. At least one pin must fail; the unmodified oracle passes them all (the pin
file itself). Every planted mistake keeps its memory accesses inside the arrays (clamped or
transposed indices only), so a mutant fails by its numbers, not by a crash.
"""
from __future__ import annotations

import ctypes as C
import importlib.util
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SOURCES = [ROOT / "oracle" / "oracle.c", ROOT / "oracle" / "render.c"]

# (name, text in oracle.c, replacement, pins that must reject it)
MUTANTS = [
    ("edge Ex averages over the Ey cells",
     "{c00, CIDX(x, y - 1, z), CIDX(x, y, z - 1), CIDX(x, y - 1, z - 1)}",
     "{c00, CIDX(x - 1, y, z), CIDX(x, y, z - 1), CIDX(x - 1, y, z - 1)}",
     ["edge_linear", "edge_isolated"]),
    ("edge Ez averages over (x, z) instead of (x, y)",
     "{c00, CIDX(x - 1, y, z), CIDX(x, y - 1, z), CIDX(x - 1, y - 1, z)}",
     "{c00, CIDX(x - 1, y, z), CIDX(x, y, z - 1), CIDX(x - 1, y, z - 1)}",
     ["edge_linear", "edge_isolated"]),
    ("edge Ey misses the diagonal neighbour",
     "{c00, CIDX(x - 1, y, z), CIDX(x, y, z - 1), CIDX(x - 1, y, z - 1)}",
     "{c00, CIDX(x - 1, y, z), CIDX(x, y, z - 1), CIDX(x, y, z - 1)}",
     ["edge_linear", "edge_isolated"]),
    ("edge neighbour mirrored at index 0 instead of clamped",
     "#define CIDX(x, y, z) IDX(g, (x) > 0 ? (x) : 0, (y) > 0 ? (y) : 0, (z) > 0 ? (z) : 0)",
     "#define CIDX(x, y, z) IDX(g, (x) >= 0 ? (x) : 1, (y) >= 0 ? (y) : 1, (z) >= 0 ? (z) : 1)",
     ["edge_linear", "edge_isolated"]),
    ("edge eps_r averages sigma",
     "ee[3 * c00 + 1] = AVG(eps, ey[0]", "ee[3 * c00 + 1] = AVG(sig, ey[0]",
     ["edge_linear", "edge_isolated"]),
    ("eAtNode x mean one cell ahead",
     "fetch_clamped(g, E, x - 1, y, z)[0]", "fetch_clamped(g, E, x + 1, y, z)[0]",
     ["node"]),
    ("eAtNode y mean reads the z component",
     "e0[1] + fetch_clamped(g, E, x, y - 1, z)[1]", "e0[1] + fetch_clamped(g, E, x, y - 1, z)[2]",
     ["node"]),
    ("bAtNode Bx mean over (x, y) instead of (y, z)",
     "fetch_clamped(g, B, x, y - 1, z - 1)[0]", "fetch_clamped(g, B, x - 1, y - 1, z)[0]",
     ["node"]),
    ("bAtNode Bz mean half a cell off in y",
     "fetch_clamped(g, B, x, y - 1, z)[2] +", "fetch_clamped(g, B, x, y + 1, z)[2] +",
     ["node"]),
    ("update_B Bx curl takes the z neighbour for d/dy",
     "float t0[3] = {ey[2] - e0[2],", "float t0[3] = {ez[2] - e0[2],",
     ["plane_wave", "energy"]),
    ("update_E Ey curl sign flipped",
     "(b0[0] - bz[0]) - (b0[2] - bx[2])", "(b0[2] - bx[2]) - (b0[0] - bz[0])",
     ["plane_wave", "energy"]),
    ("sweep half-step factor applied with the wrong sign",
     "const double hr = cos(w * 0.5 * dt), hi = sin(w * 0.5 * dt);",
     "const double hr = cos(w * 0.5 * dt), hi = -sin(w * 0.5 * dt);",
     ["sweep"]),
    # volume raymarch (oracle/render.c)
    ("texture sampled half a texel off",
     "float s = c[a] * (float)n[a] - 0.5f;", "float s = c[a] * (float)n[a];",
     ["texture"]),
    ("texture edge mirrored instead of clamped",
     "lo[a] = i < 0 ? 0 :", "lo[a] = i < 0 ? 1 :",
     ["texture"]),
    ("compositing weights by opacity instead of transmittance",
     "scl(surface, a_mat)), tr));", "scl(surface, a_mat)), alpha));",
     ["uniform", "metal"]),
    ("one raymarch step short",
     "int steps = (int)((t1 - t) / ds) + 1;", "int steps = (int)((t1 - t) / ds);",
     ["uniform", "metal"]),
    ("no early ray termination",
     "if (alpha > 0.995f) break;", "if (alpha > 2.0f) break;",
     ["metal"]),
    ("Poynting x component from the wrong pair",
     "rv3 s = v3(e.y * b.z - e.z * b.y,", "rv3 s = v3(e.y * b.y - e.z * b.z,",
     ["parallel"]),
    ("background gradient over the ray's y",
     "float h = clampf(rd.z * 0.5f + 0.5f, 0.0f, 1.0f);", "float h = clampf(rd.y * 0.5f + 0.5f, 0.0f, 1.0f);",
     ["background"]),
]


def _pins():
    spec = importlib.util.spec_from_file_location("oracle_pins_for_mutants", ROOT / "tests" / "test_oracle_pins.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    spec_r = importlib.util.spec_from_file_location("render_pins_for_mutants", ROOT / "tests" / "test_render_cpu.py")
    ren = importlib.util.module_from_spec(spec_r)
    spec_r.loader.exec_module(ren)
    return {
        "texture": [lambda O: ren.test_texture_linear_field_and_edge_clamp(O)],
        "uniform": [lambda O, m=m: ren.test_uniform_field_closed_form(O, m) for m in (0.05, 0.3, 1.5)],
        "metal": [lambda O: ren.test_metal_block_closed_form(O)],
        "parallel": [lambda O: ren.test_poynting_of_parallel_fields_is_dark(O)],
        "background": [lambda O: ren.test_empty_volume_is_background(O)],
        "edge_linear": [lambda O: mod.test_edge_average_linear_medium(O)],
        "edge_isolated": [lambda O, c=c: mod.test_edge_average_isolated_cell(O, c)
                          for c in [(4, 3, 5), (0, 3, 2), (3, 0, 0)]],
        "node": [lambda O, s=s: mod.test_node_sampling_linear_fields(O, s) for s in (0, 1)],
        "plane_wave": [lambda O: mod.test_plane_wave_exact(O)],
        "energy": [lambda O: mod.test_energy_invariant_cavity(O)],
        "sweep": [lambda O: mod.test_sweep_resistor_closed_form(O)],
    }


def _build_mutant(tmp: Path, text: str, repl: str) -> Path:
    srcs = {p.name: p.read_text() for p in SOURCES}
    sites = [n for n, t in srcs.items() if t.count(text) == 1]
    assert len(sites) == 1 and sum(t.count(text) for t in srcs.values()) == 1, \
        f"mutation site not unique or missing: {text!r}"
    srcs[sites[0]] = srcs[sites[0]].replace(text, repl)
    files = []
    for name, t in srcs.items():
        (tmp / name).write_text(t)
        files.append(str(tmp / name))
    (tmp / "oracle.h").write_text((ROOT / "oracle" / "oracle.h").read_text())
    lib = tmp / "liboracle_mut.so"
    r = subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off", "-fno-fast-math",
                        "-fopenmp", "-o", str(lib), *files, "-lm"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return lib


@pytest.mark.parametrize("name,text,repl,pins", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pins_reject_mutant(oracle, tmp_path, monkeypatch, name, text, repl, pins):
    lib_path = _build_mutant(tmp_path, text, repl)
    mlib = C.CDLL(str(lib_path))
    for fn_name, (res, args) in oracle._SIG.items():
        fn = getattr(mlib, fn_name)
        fn.restype = res
        fn.argtypes = args
    monkeypatch.setattr(oracle, "_lib", mlib)
    table = _pins()
    failures = 0
    for key in pins:
        for pin in table[key]:
            try:
                pin(oracle)
            except AssertionError:
                failures += 1
    assert failures > 0, f"no pin rejects the planted mistake: {name}"


def test_pins_accept_the_oracle(oracle):
    """The same pins, run through the same harness, pass on the unmodified oracle."""
    table = _pins()
    for key, fns in table.items():
        for pin in fns:
            pin(oracle)
