"""Batched stepper on the GPU: every scene of a batch must give the same bits as a single-grid
context running that scene alone (fields, port phasors, port series), for every launch shape,
with and without HIP-graph replay; the batched FP64 sweep must match the single-grid sweep."""
import numpy as np
import pytest

from em_lab_amd import emfdtd as E
from em_lab_amd.scene import draw_array_scene, hertzian_vacuum, spheres_dipole

from gpu_helpers import absorber_material_scene, gpu_grid, oracle_sim, rel_l2

pytestmark = pytest.mark.gpu

STEPS = 240


def _scenes():
    out = [draw_array_scene(i)[1] for i in (0, 2, 5)]
    out.append(spheres_dipole(n=48, pml=8, n_spheres=4, radius_range=(3.0, 6.0), sweep_points=16))
    out.append(hertzian_vacuum(48, 8))
    return out


def _batch(scenes, max_steps=STEPS):
    b = E.Batch(0)
    for sc in scenes:
        b.add(sc)
    b.finalize(max_steps)
    return b


def _single(sc, steps):
    g, pm, _ = gpu_grid(sc, ntff=False)
    g.step(steps)
    E_, B_ = g.fields(E.LAYOUT_SOA)
    out = dict(E=E_.ravel(), B=B_.ravel())
    if pm is not None:
        out["phasors"] = pm.phasors()
        if sc.sweep is not None:
            out["series"] = pm.read_series()
            out["sweep"] = pm.evaluate_sweep(g.dt)
    g.close()
    return out


@pytest.fixture(scope="module")
def reference_runs():
    scenes = _scenes()
    print("case: single-grid references,", [sc.size for sc in scenes], flush=True)
    return scenes, [_single(sc, STEPS) for sc in scenes]


def _check_against(b, scenes, refs, steps=STEPS):
    for i, (sc, ref) in enumerate(zip(scenes, refs)):
        Eb, Bb = b.fields(i)
        assert np.array_equal(Eb, ref["E"]), f"scene {i} E differs"
        assert np.array_equal(Bb, ref["B"]), f"scene {i} B differs"
        if "phasors" in ref:
            assert np.array_equal(b.phasors(i).ravel(), np.asarray(ref["phasors"]).ravel()), f"scene {i} phasors"
        if "series" in ref:
            s = b.series(i)
            assert np.array_equal(s.ravel(), np.asarray(ref["series"]).ravel()[:s.size]), f"scene {i} series"


def test_batch_matches_single_grids(reference_runs):
    scenes, refs = reference_runs
    print("case: batch eager, default launch", flush=True)
    b = _batch(scenes)
    info = b.info()
    assert info["scenes"] == len(scenes)
    assert info["cells"] == sum(int(np.prod(sc.size)) for sc in scenes)
    b.step(STEPS)
    _check_against(b, scenes, refs)
    # FP64 sweep on the GPU against the host sweep of the single grid
    b.evaluate_sweeps()
    for i, (sc, ref) in enumerate(zip(scenes, refs)):
        if "sweep" not in ref:
            continue
        got = b.sweep_result(i)
        want = ref["sweep"]
        assert got["residual"] == pytest.approx(want["residual"], rel=1e-12)
        for gp, wp in zip(got["ports"], want["ports"]):
            np.testing.assert_allclose(gp["points"][:, 0], wp["points"][:, 0], rtol=1e-14)
            np.testing.assert_allclose(gp["points"][:, 1:3], wp["points"][:, 1:3], rtol=1e-9, atol=1e-9)
            assert np.max(np.abs(gp["points"][:, 3] - wp["points"][:, 3])) < 1e-6
            assert gp["best"] == wp["best"]
    b.close()



def test_batch_material_in_absorber():
    """A batch holding a scene with materials in its absorber (the E pass's general path for the
    whole launch) still gives every scene the bits of its single-grid run."""
    scenes = [absorber_material_scene(), hertzian_vacuum(48, 8), draw_array_scene(0)[1]]
    print("case: batch with material in the absorber", flush=True)
    refs = [_single(sc, 120) for sc in scenes]
    b = _batch(scenes, 120)
    b.step(120)
    _check_against(b, scenes, refs, 120)
    b.close()

@pytest.mark.parametrize("shape", [(2, 0, 8), (4, 0, 1), (1, 2048, 8), (2, 1024, 1)])
def test_batch_launch_shapes(reference_runs, shape):
    scenes, refs = reference_runs
    print("case: batch launch shape", shape, flush=True)
    b = _batch(scenes)
    b.set_launch(*shape)
    assert b.launch() == shape
    b.step(STEPS)
    _check_against(b, scenes, refs)
    b.close()


def test_batch_graph_replay(reference_runs):
    """Steps in captured graphs of 32 (7 replays) plus 16 direct steps: the same bits."""
    scenes, refs = reference_runs
    print("case: batch graph replay", flush=True)
    b = _batch(scenes)
    b.set_graph(32)
    b.step(100)
    b.synchronize()
    b.step(STEPS - 100)
    assert b.step_count == STEPS
    _check_against(b, scenes, refs)
    b.close()


def test_batch_autotune_keeps_state(reference_runs):
    scenes, refs = reference_runs
    print("case: batch autotune", flush=True)
    b = _batch(scenes)
    chosen, table = b.autotune(iterations=3)
    assert len(table) == 12 and np.all(table[:, 2] > 0)
    assert b.launch() == chosen
    for i in range(len(scenes)):
        Eb, Bb = b.fields(i)
        assert not np.any(Eb) and not np.any(Bb)
    b.step(STEPS)
    _check_against(b, scenes, refs)
    with pytest.raises(E.EmfdError):
        b.autotune(1)  # only before the first step
    b.close()


def test_batch_random_fields_and_reset():
    """Non-zero fields everywhere (absorber, wires, feed) through the batch and a single grid."""
    scenes = [draw_array_scene(7)[1], draw_array_scene(11)[1]]
    print("case: batch random fields", [sc.size for sc in scenes], flush=True)
    b = _batch(scenes, max_steps=64)
    rng = np.random.default_rng(5)
    init = []
    for i, sc in enumerate(scenes):
        n = int(np.prod(sc.size))
        e0 = rng.standard_normal(3 * n).astype(np.float32)
        b0 = rng.standard_normal(3 * n).astype(np.float32)
        init.append((e0, b0))
        b.set_fields(i, e0, b0)
    b.step(20)
    for i, sc in enumerate(scenes):
        g, _, _ = gpu_grid(sc, ntff=False)
        g.set_fields(init[i][0].reshape(3, -1), init[i][1].reshape(3, -1), E.LAYOUT_SOA)
        g.step(20)
        Eg, Bg = g.fields(E.LAYOUT_SOA)
        Eb, Bb = b.fields(i)
        assert np.array_equal(Eb, Eg.ravel()) and np.array_equal(Bb, Bg.ravel())
        g.close()
    b.reset()
    assert b.step_count == 0
    Eb, Bb = b.fields(0)
    assert not np.any(Eb) and not np.any(Bb)
    with pytest.raises(E.EmfdError):
        b.step(65)  # beyond max_steps
    b.close()


def test_batch_scene_against_oracle():
    """One batched antenna scene against the CPU oracle directly."""
    _, sc = draw_array_scene(3)
    print("case: batch vs oracle", sc.size, flush=True)
    b = _batch([sc], max_steps=120)
    b.step(120)
    sim = oracle_sim(sc, ntff=False)
    sim.step(120)
    Eb, Bb = b.fields(0)
    n = int(np.prod(sc.size))
    Eo = sim.E[:, :3].T.ravel()
    Bo = sim.B[:, :3].T.ravel()
    assert rel_l2(Eb, Eo) <= 1e-5 and rel_l2(Bb, Bo) <= 1e-5
    assert Eb.size == 3 * n
    b.close()
