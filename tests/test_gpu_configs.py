"""The workloads at their stated sizes on the GPU:

- the 256^3 sphere + thin-wire dipole scene with every monitor on (ports with a 32-point sweep
  recording, far field at three wavelengths, a flux box, probes) against the oracle;
- the precision form: the 192^3 copy of the 768^3 precision scene for 4000 steps in FP32,
  binary16 and bfloat16 against the FP64 stepper, held to the error table the accuracy study
  reports; the 768^3 scene itself for 4000 steps in FP32 and binary16;
- the batched form: 512 seeded thin-wire antenna scenes in one batch for 4000 steps with the
  64-bin sweep, four of them against single-grid runs of the same scene bit for bit;
- the learned form: the committed surrogate on 65,536 seeded candidates (its output on 512
  sampled rows against the CPU restatement of the network), screened down to 256 designs that
  the batched FDTD verifies."""
import math

import numpy as np
import pytest

from em_lab_amd import emfdtd as E
from em_lab_amd import precision as PR
from em_lab_amd import surrogate as S
from em_lab_amd.scene import draw_array_scene, spheres_array, spheres_dipole
from oracle import mlp as M

from gpu_helpers import gpu_grid, oracle_sim, rel_l2

pytestmark = pytest.mark.gpu


def test_monitors_256_against_oracle():
    """256^3, 10-cell absorber, 32 spheres, thin-wire dipole on a 50 ohm feed, every monitor on:
    after 40 steps the fields, port phasors and series and the far-field phasors of three
    wavelengths equal the oracle's bit for bit. The GPU then runs on to 460 steps (the pulse has
    crossed the Huygens box), and its pattern on 400 directions, radiated power, flux box and
    probes are held against the oracle's evaluation of the same phasors and fields."""
    print("case: 256^3 all monitors, 40 steps against the oracle, 460 on the GPU", flush=True)
    sc = spheres_dipole(n=256, pml=10)
    g, pm, ff = gpu_grid(sc)
    wls = [sc.far_field.wavelength, sc.far_field.wavelength * 0.75, sc.far_field.wavelength * 1.5]
    ff.set_wavelengths(wls)
    steps = 40
    sim = oracle_sim(sc, gpu_ntff_setup=ff.setup, capacity=steps)
    extra = [sim.add_ntff_wavelength(w) for w in wls[1:]]
    g.step(steps)
    sim.step(steps)
    Eg, Bg = g.fields(E.LAYOUT_TEXEL)
    got = dict(E=Eg[:, :3], B=Bg[:, :3], port=pm.phasors(), series=pm.read_series(), ff0=ff.phasors(0),
               ff1=ff.phasors(1), ff2=ff.phasors(2))
    want = dict(E=sim.E[:, :3], B=sim.B[:, :3], port=sim.ports["phasors"][:1], series=sim.ports["series"][:steps],
                ff0=sim.ntff["phasors"], ff1=extra[0], ff2=extra[1])
    assert np.abs(Eg[:, :3]).max() > 0
    for key in got:
        r = rel_l2(got[key], want[key])
        print(f"{key}: rel L2 {r:.2e}, bit-identical {np.array_equal(got[key], want[key])}", flush=True)
        assert r <= 1e-5, key
    del got, want
    g.step(420)
    Eg, Bg = g.fields(E.LAYOUT_TEXEL)
    sim.E[:] = Eg
    sim.B[:] = Bg
    del Eg, Bg
    sim.ntff["phasors"][:] = ff.phasors(0)
    extra[0][:] = ff.phasors(1)
    extra[1][:] = ff.phasors(2)
    rng = np.random.default_rng(3)
    dirs = rng.normal(size=(400, 3))
    dirs = np.ascontiguousarray(np.concatenate([dirs / np.linalg.norm(dirs, axis=1, keepdims=True),
                                                np.zeros((400, 1))], 1), np.float32)
    for q in range(3):
        f2g, Pg = ff.radiation(dirs, q)
        f2o, Po = sim.ntff_eval_f32(dirs, extra=None if q == 0 else q - 1)
        assert Po > 0
        assert Pg == pytest.approx(Po, rel=1e-4)
        k = 2 * math.pi / wls[q]
        Dg = 10 * np.log10(k * k * f2g / (8 * math.pi * Pg))
        Do = 10 * np.log10(k * k * f2o / (8 * math.pi * Po))
        strong = Do > Do.max() - 30
        print(f"pattern lambda {wls[q]}: P {Pg:.4e} vs {Po:.4e}, max |dD| {np.max(np.abs(Dg[strong] - Do[strong])):.2e} dB",
              flush=True)
        assert np.max(np.abs(Dg[strong] - Do[strong])) <= 0.05
    an = E.FieldAnalyzer(g)
    fl = an.measure_flux(sc.fluxes)[0]
    fo = sim.box_flux(sc.fluxes[0].box_min, sc.fluxes[0].box_max)
    print(f"flux box {fl:.6e} vs {fo:.6e}", flush=True)
    assert fo != 0.0
    assert fl == pytest.approx(fo, rel=1e-4)
    for p in [(128.0, 128.0, 128.5), (100.3, 140.7, 120.2)]:
        s = an.probe(p)[0]
        o = sim.probe(p)
        np.testing.assert_allclose(s["E"], o[0:3], rtol=1e-4, atol=1e-9)
        np.testing.assert_allclose(s["B"], o[3:6], rtol=1e-4, atol=1e-9)
    g.close()


# The accuracy study's figures with a margin (192^3 copy of the precision scene, 4000 steps,
# against the FP64 stepper): S11 over 16 ports x 32 bins floored at -40 dB, directivity within
# 30 dB of the peak, radiated power.
PRECISION_BOUNDS = {
    "f32": dict(e500=1e-5, s11_max=1e-4, s11_median=1e-5, d_max=0.01, d_peak=0.001, power=1e-4),
    "f16": dict(e500=3e-2, s11_max=0.02, s11_median=0.002, d_max=0.2, d_peak=0.01, power=2e-4),
    "bf16": dict(e500=0.2, s11_max=0.05, s11_median=0.01, d_max=2.5, d_peak=0.06, power=3e-2),
}


@pytest.fixture(scope="module")
def precision_report():
    print("case: precision study 192^3, 4000 steps, f64 / f32 / f16 / bf16", flush=True)
    return PR.study(192, 4000, log=lambda m: print(m, flush=True))


def test_precision_study_192(precision_report):
    for name, b in PRECISION_BOUNDS.items():
        r = precision_report["runs"][name]
        assert r["field_rel_l2"]["500"]["E"] <= b["e500"], (name, r)
        assert r["s11_max_abs_err_db"] <= b["s11_max"], (name, r)
        assert r["s11_median_abs_err_db"] <= b["s11_median"], (name, r)
        assert r["directivity_max_abs_err_db"] <= b["d_max"], (name, r)
        assert abs(r["max_directivity_err_db"]) <= b["d_peak"], (name, r)
        assert r["radiated_power_rel_err"] <= b["power"], (name, r)


def test_precision_scene_768(precision_report):
    """The 768^3 precision scene for 4000 steps, FP32 and binary16 (the scales the 192^3 study
    took from its FP64 peaks): finite fields, rung down ports, S11 of every port within the
    binary16 bound of the FP32 run, peak directivity within 0.01 dB."""
    print("case: precision scene 768^3, 4000 steps, f32 and f16", flush=True)
    sc = spheres_array(768)
    k = precision_report["scale_log2_f16"]
    out = {}
    for name, st in (("f32", E.STORAGE_F32), ("f16", E.STORAGE_F16)):
        g, pm, ff = PR._setup(sc, st, k if st == E.STORAGE_F16 else None)
        g.step(4000)
        sw = pm.evaluate_sweep(g.dt)
        res = ff.evaluate()
        Ef, _ = g.fields(E.LAYOUT_SOA)
        assert np.all(np.isfinite(Ef)), name
        del Ef
        assert sw["residual"] < 1e-2, (name, sw["residual"])
        out[name] = (np.stack([p["points"][:, 3] for p in sw["ports"]]), res["max_directivity_dbi"])
        print(f"768^3 {name}: residual {sw['residual']:.2e}, max directivity {res['max_directivity_dbi']:.4f} dBi",
              flush=True)
        g.close()
    s32 = np.maximum(out["f32"][0], PR.S11_FLOOR_DB)
    s16 = np.maximum(out["f16"][0], PR.S11_FLOOR_DB)
    print("768^3 S11 f16 vs f32: max", np.abs(s16 - s32).max(), "dB", flush=True)
    assert s32.shape == (16, 32)
    assert np.abs(s16 - s32).max() <= PRECISION_BOUNDS["f16"]["s11_max"]
    assert abs(out["f16"][1] - out["f32"][1]) <= PRECISION_BOUNDS["f16"]["d_peak"]


def test_batch_512_scenes():
    """512 scenes of the thin-wire array generator (48x48x96 .. 128x128x160), 4000 steps, 64-bin
    sweep on the GPU: every scene finite and rung down; scenes 0, 171, 342 and 511 equal their
    single-grid runs bit for bit (fields, phasors, series) and in the sweep to 1e-9."""
    print("case: batch of 512 array scenes, 4000 steps", flush=True)
    steps = 4000
    scenes = [draw_array_scene(i)[1] for i in range(512)]
    b = E.Batch(0)
    for sc in scenes:
        b.add(sc)
    b.finalize(steps)
    b.autotune(3)
    b.step(steps)
    b.evaluate_sweeps()
    res = np.array([b.sweep_result(i)["residual"] for i in range(512)])
    print("residual: median", np.median(res), "max", res.max(), flush=True)
    assert np.all(np.isfinite(res)) and np.all(res < 1e-2)
    for i in (0, 171, 342, 511):
        sc = scenes[i]
        Eb, Bb = b.fields(i)
        assert np.all(np.isfinite(Eb))
        g, pm, _ = gpu_grid(sc, ntff=False)
        g.step(steps)
        Eg, Bg = g.fields(E.LAYOUT_SOA)
        assert np.array_equal(Eb, Eg.ravel()) and np.array_equal(Bb, Bg.ravel()), i
        assert np.array_equal(b.phasors(i).ravel(), pm.phasors().ravel()), i
        sb = b.series(i)
        assert np.array_equal(sb.ravel(), pm.read_series().ravel()[:sb.size]), i
        # the batched FP64 sweep on the GPU against the host sweep of the single grid
        sw = pm.evaluate_sweep(g.dt)
        pb = b.sweep_result(i)["ports"][0]["points"]
        ps = sw["ports"][0]["points"]
        assert pb.shape == (64, 5)
        np.testing.assert_allclose(pb[:, 0], ps[:, 0], rtol=1e-14)
        np.testing.assert_allclose(pb[:, 1:3], ps[:, 1:3], rtol=1e-9, atol=1e-9)
        assert np.max(np.abs(pb[:, 3] - ps[:, 3])) < 1e-6
        g.close()
    b.close()


def test_surrogate_65536_screen_to_256():
    """The committed surrogate on 65,536 seeded candidates: 512 sampled rows against the CPU
    restatement of the network (bfloat16 rounding where the device rounds), the best 256 by
    predicted -10 dB bandwidth verified by the batched FDTD with 256 random candidates beside
    them; the picks' verified bandwidth beats the random designs'."""
    print("case: surrogate 65,536 candidates, 512 rows vs oracle, top 256 verified", flush=True)
    n = 65536
    geoms, sizes, X = S.candidates(n)
    m = S.load_surrogate(max_batch=n)
    members = m.members if isinstance(m, S.Ensemble) else [m]
    y = m.forward(X)
    assert y.shape == (n, 128) and np.all(np.isfinite(y))
    rows = np.random.default_rng(5).choice(n, 512, replace=False)
    yo = np.mean([M.forward(mm.params(), S.DIMS, X[rows], emulate_bf16=True)[0] for mm in members], axis=0)
    err = rel_l2(y[rows], yo)
    print("forward vs oracle on 512 rows: rel L2", err, flush=True)
    assert err <= 2e-3
    idx, s_pred, bw_pred = S.screen(m, X, 256)
    m.close()
    assert len(set(idx.tolist())) == 256
    rand = np.random.default_rng(1).choice(np.setdiff1d(np.arange(n), idx), 256, replace=False)
    both = np.concatenate([idx, rand])
    s_fdtd, _, res = S.verify(geoms, sizes, both)
    assert np.all(res < 1e-2)
    bw = S.bandwidth_10db(s_fdtd, S.sweep_nu())
    err_db = S.s11_error_db(s_pred[idx], s_fdtd[:256])
    precision = float(np.mean(err_db <= 0.5))
    print(f"verified bandwidth: picks {bw[:256].mean():.4f}, random {bw[256:].mean():.4f}; "
          f"pick precision (S11 within 0.5 dB at every bin) {precision:.3f}", flush=True)
    assert bw[:256].mean() > 1.5 * bw[256:].mean()
    assert precision >= 0.5
