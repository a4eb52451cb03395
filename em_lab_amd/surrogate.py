"""Learned screening of thin-wire antenna designs: batched-FDTD training data, the MLP surrogate
(18 geometry features -> 64-bin complex port response), candidate screening and FDTD
verification.

The network's 128 outputs are the real and imaginary parts of the port reflection coefficient
Gamma = (Z - R) / (Z + R), R = 50 ohm, at the 64 sweep bins; the impedance follows as
Z = R (1 + Gamma) / (1 - Gamma). Gamma is the bounded Moebius image of Z, so a mean-squared error
on it weighs every design and bin alike (the raw reactance of a short wire spans kilo-ohms).

Data come from the batched stepper (emfd_batch_*) on the scene generator
em_lab_amd.scene.draw_array_scene: seeded, one xoshiro256** stream per scene index.
"""
from __future__ import annotations

import argparse
import math
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from em_lab_amd.scene import (ARRAY_LAMBDA0, ARRAY_MAX_ELEMENTS, ArrayGeometry, DEFAULT_SEED,  # noqa: E402
                              array_scene, draw_array_design, draw_array_designs, draw_array_scene)

N_FEATURES = 3 * ARRAY_MAX_ELEMENTS
N_BINS = 64
DIMS = (N_FEATURES, 256, 256, 256, 2 * N_BINS)
R0 = 50.0
STEPS = 4000
CANDIDATE_SEED = 7777
S11_FLOOR_DB = -25.0   # S11 below this counts as matched; differences are taken above it
WEIGHTS = Path(__file__).resolve().parent / "data" / "surrogate_weights.npz"


def features(geom: ArrayGeometry, lambda0: float = ARRAY_LAMBDA0) -> np.ndarray:
    """18 inputs: per element (length / lambda0, spacing / lambda0, log10(radius / lambda0) + 3.5),
    zeros for absent elements (every present element has a non-zero length entry: the mask)."""
    f = np.zeros(N_FEATURES, np.float32)
    for i in range(geom.elements):
        f[3 * i + 0] = geom.lengths[i] / lambda0
        f[3 * i + 1] = geom.spacings[i] / lambda0
        f[3 * i + 2] = math.log10(geom.radii[i] / lambda0) + 3.5
    return f


def features_many(geoms) -> np.ndarray:
    """features(g) for every geometry, as one (designs, 18) array (the same float32 values)."""
    m = ARRAY_MAX_ELEMENTS
    count = len(geoms)
    n = np.array([g.elements for g in geoms], np.int64)
    lengths = np.zeros((count, m), np.float64)
    spacings = np.zeros((count, m), np.float64)
    logs = np.zeros((count, m), np.float64)
    present = np.arange(m)[None, :] < n[:, None]
    rows = np.nonzero(present)
    lengths[rows] = [v for g in geoms for v in g.lengths]
    spacings[rows] = [v for g in geoms for v in g.spacings]
    # log10 through math.log10, as features() takes it
    logs[rows] = [math.log10(r / ARRAY_LAMBDA0) + 3.5 for g in geoms for r in g.radii]
    f = np.zeros((count, N_FEATURES), np.float32)
    f[:, 0::3] = np.where(present, lengths / ARRAY_LAMBDA0, 0.0)
    f[:, 1::3] = np.where(present, spacings / ARRAY_LAMBDA0, 0.0)
    f[:, 2::3] = np.where(present, logs, 0.0)
    return f


def gamma_to_z(g: np.ndarray) -> np.ndarray:
    return R0 * (1.0 + g) / (1.0 - g)


def z_to_gamma(z: np.ndarray) -> np.ndarray:
    return (z - R0) / (z + R0)


def s11_db(g: np.ndarray) -> np.ndarray:
    return 20.0 * np.log10(np.maximum(np.abs(g), 1e-12))


def bandwidth_10db(s11: np.ndarray, nu: np.ndarray) -> np.ndarray:
    """Width in normalised frequency of the widest run of bins with S11 < -10 dB (0 if none);
    s11: (designs, bins)."""
    below = s11 < -10.0
    run = np.zeros(s11.shape[0], np.int64)
    best = np.zeros(s11.shape[0], np.int64)
    for b in range(s11.shape[1]):
        run = (run + 1) * below[:, b]
        best = np.maximum(best, run)
    return best * (nu[1] - nu[0])


def targets_from_z(z: np.ndarray, power: float = 0.0) -> np.ndarray:
    """(designs, 64) complex Z -> (designs, 128) network targets: Re then Im of the companded
    reflection coefficient Gamma |Gamma|^-power (power 0: Gamma itself). With power > 0 the
    squared error weighs a deviation at small |Gamma| (a deep resonance) more, closer to the
    screening test, which is in dB: a deviation d at |Gamma| costs about 8.69 d / |Gamma| dB."""
    g = z_to_gamma(z)
    if power:
        a = np.abs(g)
        g = np.where(a > 0, g * np.power(np.maximum(a, 1e-300), -power), 0)
    return np.concatenate([g.real, g.imag], axis=1).astype(np.float32)


def gamma_from_output(y: np.ndarray, power: float = 0.0) -> np.ndarray:
    """Inverse of targets_from_z: |output| = |Gamma|^(1 - power), same phase."""
    o = y[:, :N_BINS].astype(np.float64) + 1j * y[:, N_BINS:].astype(np.float64)
    if power:
        a = np.abs(o)
        o = o * np.power(np.maximum(a, 1e-300), power / (1.0 - power))
    return o


def output_power(m) -> float:
    """The target companding power a loaded or trained surrogate was trained with."""
    return float(getattr(m, "target_power", 0.0))


def s11_error_db(s11_a: np.ndarray, s11_b: np.ndarray) -> np.ndarray:
    """Per design: max over bins of |S11_a - S11_b| in dB, both floored at S11_FLOOR_DB."""
    a = np.maximum(s11_a, S11_FLOOR_DB)
    b = np.maximum(s11_b, S11_FLOOR_DB)
    return np.max(np.abs(a - b), axis=1)


def sweep_nu(lambda0: float = ARRAY_LAMBDA0, points: int = N_BINS) -> np.ndarray:
    """Normalised frequency f / f0 of the sweep bins (0.5 .. 1.5)."""
    lo, hi = 1.0 / (lambda0 / 0.5), 1.0 / (lambda0 / 1.5)
    return (lo + (hi - lo) * np.arange(points) / (points - 1)) * lambda0


def run_fdtd(make_scene, n: int, steps: int = STEPS, batch_size: int = 1024, log=None):
    """Batched FDTD of scenes make_scene(0) .. make_scene(n-1): complex Z (designs, 64) at the
    port of element 0 and the ring-down residual. Scenes are built and run in chunks of
    batch_size so that only one chunk's host arrays exist at a time."""
    from em_lab_amd import emfdtd as E
    zs, res = [], []
    for c0 in range(0, n, batch_size):
        c1 = min(c0 + batch_size, n)
        t0 = time.time()
        b = E.Batch(0)
        for i in range(c0, c1):
            b.add(make_scene(i))
        b.finalize(steps)
        b.autotune(3)
        t1 = time.time()
        b.step(steps)
        b.evaluate_sweeps()
        for i in range(c1 - c0):
            r = b.sweep_result(i)
            pts = r["ports"][0]["points"]
            zs.append(pts[:, 1] + 1j * pts[:, 2])
            res.append(r["residual"])
        b.close()
        if log:
            log(f"fdtd designs {c0}..{c1}: build {t1 - t0:.1f} s, {steps} steps + sweep {time.time() - t1:.1f} s")
    return np.array(zs), np.array(res)


def generate(start: int, count: int, out: Path, seed: int = DEFAULT_SEED, batch_size: int = 1024):
    """Training data: scenes start .. start+count-1 of the seeded generator through batched FDTD."""
    geoms, sizes = [], []

    def make(i):
        g, sc = draw_array_scene(start + i, seed)
        geoms.append(g)
        sizes.append(sc.size)
        return sc

    z, res = run_fdtd(make, count, STEPS, batch_size, log=lambda m: print(m, flush=True))
    X = np.stack([features(g) for g in geoms])
    np.savez_compressed(out, index=np.arange(start, start + count), X=X, z=z, residual=res, size=np.array(sizes))
    return X, z


def load_data(paths):
    X, Z, R = [], [], []
    for p in paths:
        d = np.load(p)
        X.append(d["X"])
        Z.append(d["z"])
        R.append(d["residual"])
    return np.concatenate(X), np.concatenate(Z), np.concatenate(R)


def train_surrogate(X, Z, epochs=400, batch=256, lr=(2e-3, 2e-5), seed=1, holdout=1024, log=None, power=0.0):
    """Trains the surrogate on the GPU (emfd_mlp_*) with the last `holdout` designs held out, on
    targets companded with `power` (targets_from_z). Returns (Mlp, report dict)."""
    from em_lab_amd import emfdtd as E
    T = targets_from_z(Z, power)
    ntr = X.shape[0] - holdout
    m = E.Mlp(DIMS, max_batch=65536)
    m.init_params(seed=seed)
    m.set_adam(0.9, 0.999, 1e-8, 0.0)
    m.target_power = float(power)
    m.set_dataset(X[:ntr], T[:ntr])
    t0 = time.time()
    losses = m.train(batch=batch, epochs=epochs, lr_start=lr[0], lr_end=lr[1], seed=seed)
    t_train = time.time() - t0
    rep = dict(train_loss=float(losses[-1]), first_loss=float(losses[0]), epochs=epochs, train_s=t_train,
               train_designs=int(ntr))
    if holdout:
        y = m.forward(X[ntr:])
        s_pred = s11_db(gamma_from_output(y, power))
        s_true = s11_db(z_to_gamma(Z[ntr:]))
        err = s11_error_db(s_pred, s_true)
        rep.update(holdout=int(holdout), holdout_mse=float(np.mean((y - T[ntr:]) ** 2)),
                   holdout_s11_err_median_db=float(np.median(err)), holdout_s11_err_p90_db=float(np.percentile(err, 90)),
                   holdout_within_0p5db=float(np.mean(err <= 0.5)))
    if log:
        log(f"trained {epochs} epochs in {t_train:.1f} s: {rep}")
    return m, rep


class Ensemble:
    """Several surrogate networks (same widths, different seeds) whose outputs are averaged."""

    def __init__(self, members):
        self.members = list(members)
        self.dims = self.members[0].dims
        self.target_power = output_power(self.members[0])

    def forward(self, X):
        y = self.members[0].forward(X).astype(np.float64)
        for m in self.members[1:]:
            y += m.forward(X)
        return (y / len(self.members)).astype(np.float32)

    def time_forward(self, batch, iterations=20):
        return sum(m.time_forward(batch, iterations) for m in self.members)

    def params(self):
        return np.stack([m.params() for m in self.members])

    def close(self):
        for m in self.members:
            m.close()


def train_ensemble(X, Z, members=4, epochs=600, batch=256, lr=(2e-3, 2e-5), holdout=1024, log=None, power=0.0):
    """`members` networks trained from different seeds on the same data; the report scores their
    averaged prediction on the held-out designs."""
    nets, reps = [], []
    for k in range(members):
        m, rep = train_surrogate(X, Z, epochs=epochs, batch=batch, lr=lr, seed=1 + k, holdout=holdout, log=log,
                                 power=power)
        nets.append(m)
        reps.append(rep)
    ens = Ensemble(nets)
    rep = dict(members=members, member_holdout_s11_err_median_db=[r.get("holdout_s11_err_median_db") for r in reps],
               train_s=round(sum(r["train_s"] for r in reps), 2))
    if holdout:
        ntr = X.shape[0] - holdout
        y = ens.forward(X[ntr:])
        err = s11_error_db(s11_db(gamma_from_output(y, power)), s11_db(z_to_gamma(Z[ntr:])))
        rep.update(holdout_s11_err_median_db=float(np.median(err)), holdout_s11_err_p90_db=float(np.percentile(err, 90)),
                   holdout_within_0p5db=float(np.mean(err <= 0.5)))
    if log:
        log(f"ensemble: {rep}")
    return ens, rep


def save_weights(m, path: Path = WEIGHTS, report=None) -> None:
    """Parameters of one network (params: n) or of an ensemble (params: members x n)."""
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(path, params=m.params(), dims=np.array(m.dims), target_power=np.array(output_power(m)),
                        report=np.array(repr(report or {})))


def load_surrogate(path: Path = WEIGHTS, max_batch: int = 65536):
    from em_lab_amd import emfdtd as E
    d = np.load(path)
    dims = tuple(int(v) for v in d["dims"])
    params = np.atleast_2d(d["params"])
    power = float(d["target_power"]) if "target_power" in d.files else 0.0
    nets = []
    for p in params:
        m = E.Mlp(dims, max_batch=max_batch)
        m.set_params(p)
        m.target_power = power
        nets.append(m)
    return nets[0] if len(nets) == 1 else Ensemble(nets)


def candidates(n: int, seed: int = CANDIDATE_SEED):
    """n seeded candidate designs: (geometries, sizes, features), drawn with all generator
    streams stepped together (draw_array_designs gives draw_array_design's designs)."""
    geoms, sizes = draw_array_designs(0, n, seed)
    return geoms, sizes, features_many(geoms)


def screen(m, X: np.ndarray, top: int = 256):
    """Surrogate S11 of every candidate and the `top` by predicted -10 dB bandwidth (ties broken
    by the lowest predicted S11)."""
    y = m.forward(X)
    s = s11_db(gamma_from_output(y, output_power(m)))
    bw = bandwidth_10db(s, sweep_nu())
    order = np.lexsort((s.min(axis=1), -bw))
    return order[:top], s, bw


def verify(geoms, sizes, idx, steps: int = STEPS, batch_size: int = 1024):
    """Batched FDTD of the chosen candidates: S11 (designs, 64) and Z."""
    z, res = run_fdtd(lambda i: array_scene(geoms[idx[i]], sizes[idx[i]]), len(idx), steps, batch_size)
    return s11_db(z_to_gamma(z)), z, res


HARD_SEED = 4242  # candidate pool for hard examples: a generator seed other than the screening's


def mine_hard(count: int, pool: int, out: Path, seed: int = HARD_SEED, batch_size: int = 1024, log=None):
    """Hard examples for the surrogate: of `pool` candidates of generator seed `seed` (not the
    screening benchmark's CANDIDATE_SEED, so its candidates stay unseen), the `count` the current
    surrogate ranks best by predicted -10 dB bandwidth, i.e. the extreme designs screening picks
    and where the surrogate errs most, simulated by batched FDTD and saved in the training-data
    format (index = position in the pool)."""
    m = load_surrogate()
    geoms, sizes = draw_array_designs(0, pool, seed)
    X = features_many(geoms)
    chunk = 65536
    y = np.concatenate([m.forward(X[i:i + chunk]) for i in range(0, pool, chunk)])
    power = output_power(m)
    m.close()
    s = s11_db(gamma_from_output(y, power))
    bw = bandwidth_10db(s, sweep_nu())
    idx = np.lexsort((s.min(axis=1), -bw))[:count]
    z, res = run_fdtd(lambda i: array_scene(geoms[idx[i]], sizes[idx[i]]), len(idx), STEPS, batch_size, log=log)
    np.savez_compressed(out, index=idx, seed=np.array(seed), X=X[idx], z=z, residual=res,
                        size=np.array([sizes[i] for i in idx]))
    return X[idx], z


def plain_fdtd_rate(n: int = 4, seed: int = CANDIDATE_SEED, steps: int = STEPS) -> float:
    """Designs per second of one-at-a-time FDTD (a single-grid context per design, its sweep)."""
    from em_lab_amd import emfdtd as E
    t = time.time()
    for i in range(n):
        g, size = draw_array_design(i, seed)
        sc = array_scene(g, size)
        grid = E.SimulationGrid(0)
        grid.configure(sc.grid_config())
        s, e = sc.flat_materials()
        grid.upload_materials(s, e, sc.subcell)
        grid.set_sources(sc.sources)
        pm = E.PortMonitor(grid)
        pm.configure(sc.ports, 4.0, sc.sweep, sc.cell_size)
        grid.step(steps)
        pm.evaluate_sweep(grid.dt)
        grid.close()
    return n / (time.time() - t)


def screening_run(n_candidates: int = 65536, top: int = 256, plain_designs: int = 4, log=None,
                  weights: Path = WEIGHTS):
    """The learned screening pipeline end to end: candidates -> surrogate -> top designs ->
    batched FDTD verification. Returns a report with designs/s against one-at-a-time FDTD."""
    t0 = time.time()
    geoms, sizes, X = candidates(n_candidates)
    t1 = time.time()
    m = load_surrogate(weights)
    t2 = time.time()
    idx, s_pred, bw_pred = screen(m, X, top)
    t3 = time.time()
    s_fdtd, _, _ = verify(geoms, sizes, idx)
    t4 = time.time()
    err = s11_error_db(s_pred[idx], s_fdtd)
    nu = sweep_nu()
    bw_fdtd = bandwidth_10db(s_fdtd, nu)
    ms_fwd = m.time_forward(min(n_candidates, 65536), 10)
    members = len(m.members) if isinstance(m, Ensemble) else 1
    m.close()
    plain = plain_fdtd_rate(plain_designs)
    screen_s = (t1 - t0) + (t3 - t2) + (t4 - t3)
    rep = dict(candidates=n_candidates, verified=int(top), candidate_gen_s=round(t1 - t0, 3),
               inference_s=round(t3 - t2, 3), verify_s=round(t4 - t3, 3),
               designs_per_s=round(n_candidates / screen_s, 2), plain_fdtd_designs_per_s=round(plain, 3),
               speedup=round(n_candidates / screen_s / plain, 1),
               pick_precision=round(float(np.mean(err <= 0.5)), 4), picks_within_0p5db=int(np.sum(err <= 0.5)),
               s11_err_median_db=round(float(np.median(err)), 3),
               s11_err_max_db=round(float(np.max(err)), 3),
               verified_bw_mean=round(float(np.mean(bw_fdtd)), 4), predicted_bw_mean=round(float(np.mean(bw_pred[idx])), 4),
               surrogate_members=members, forward_ms_65536=round(ms_fwd, 4),
               forward_tflops=round(members * 2 * 65536 * sum(i * o for i, o in zip(DIMS[:-1], DIMS[1:])) /
                                    (ms_fwd * 1e-3) / 1e12, 2))
    if log:
        log(repr(rep))
    return rep


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate", help="batched-FDTD training data")
    g.add_argument("--start", type=int, default=0)
    g.add_argument("--count", type=int, default=1024)
    g.add_argument("--batch", type=int, default=1024)
    g.add_argument("--out", type=Path, required=True)
    tr = sub.add_parser("train", help="train the surrogate on generated data and save its weights")
    tr.add_argument("--data", type=Path, nargs="+", required=True)
    tr.add_argument("--epochs", type=int, default=400)
    tr.add_argument("--batch", type=int, default=256)
    tr.add_argument("--members", type=int, default=1, help="ensemble size (networks averaged)")
    tr.add_argument("--out", type=Path, default=WEIGHTS)
    tr.add_argument("--target-power", type=float, default=0.0,
                    help="train on Gamma |Gamma|^-p (0: Gamma itself; see targets_from_z)")
    mh = sub.add_parser("mine", help="hard examples: the best-ranked designs of a separate candidate pool, by FDTD")
    mh.add_argument("--count", type=int, default=2048)
    mh.add_argument("--pool", type=int, default=262144)
    mh.add_argument("--seed", type=int, default=HARD_SEED)
    mh.add_argument("--out", type=Path, required=True)
    sc = sub.add_parser("screen", help="screen seeded candidates and verify the best by batched FDTD")
    sc.add_argument("--candidates", type=int, default=65536)
    sc.add_argument("--top", type=int, default=256)
    sc.add_argument("--weights", type=Path, default=WEIGHTS)
    fw = sub.add_parser("forward", help="inference only: the surrogate on seeded candidates, timed")
    fw.add_argument("--candidates", type=int, default=65536)
    fw.add_argument("--iterations", type=int, default=10)
    fw.add_argument("--weights", type=Path, default=WEIGHTS)
    args = ap.parse_args(argv)
    if args.cmd == "forward":
        t = time.time()
        _, _, X = candidates(args.candidates)
        print(f"{args.candidates} candidates in {time.time() - t:.1f} s", flush=True)
        m = load_surrogate(args.weights, max_batch=args.candidates)
        members = len(m.members) if isinstance(m, Ensemble) else 1
        flop = members * 2 * args.candidates * sum(i * o for i, o in zip(DIMS[:-1], DIMS[1:]))
        for it in range(args.iterations):
            t = time.time()
            m.forward(X)
            dt = time.time() - t
            print(f"forward {it}: {dt * 1e3:.2f} ms ({flop / dt / 1e12:.1f} TFLOP/s with the host copies)", flush=True)
        m.close()
        return 0
    if args.cmd == "train":
        X, Z, _ = load_data(args.data)
        if args.members > 1:
            m, rep = train_ensemble(X, Z, members=args.members, epochs=args.epochs, batch=args.batch,
                                    log=lambda s: print(s, flush=True), power=args.target_power)
        else:
            m, rep = train_surrogate(X, Z, epochs=args.epochs, batch=args.batch, log=lambda s: print(s, flush=True),
                                     power=args.target_power)
        save_weights(m, args.out, rep)
        print(f"saved {args.out}", flush=True)
        return 0
    if args.cmd == "screen":
        screening_run(args.candidates, args.top, log=lambda s: print(s, flush=True), weights=args.weights)
        return 0
    if args.cmd == "mine":
        t = time.time()
        mine_hard(args.count, args.pool, args.out, seed=args.seed, log=lambda m: print(m, flush=True))
        print(f"mined {args.count} of {args.pool} designs in {time.time() - t:.1f} s -> {args.out}", flush=True)
        return 0
    if args.cmd == "generate":
        t = time.time()
        generate(args.start, args.count, args.out, batch_size=args.batch)
        print(f"generated {args.count} designs in {time.time() - t:.1f} s -> {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
