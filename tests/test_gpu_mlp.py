"""Surrogate MLP kernels on the GPU (BF16 MFMA, FP32 accumulation) against the CPU FP64
restatement: forward outputs, loss, and every parameter gradient (the FP64 gradient check),
with the oracle rounding to bfloat16 where the device does (tight) and without (loose: the
BF16 error budget). Training must reduce the loss."""
import numpy as np
import pytest

from em_lab_amd import emfdtd as E
from oracle import mlp as M

from gpu_helpers import rel_l2

pytestmark = pytest.mark.gpu

DIMS = (18, 256, 256, 256, 128)


def _inputs(rows, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 1.5, (rows, DIMS[0])).astype(np.float32)
    x[:, 9:] *= rng.integers(0, 2, (rows, 1))  # absent elements as zero padding
    t = rng.normal(0.0, 0.5, (rows, DIMS[-1])).astype(np.float32)
    return x, t


@pytest.mark.parametrize("rows", [1, 100, 300])
def test_forward_matches_oracle(rows):
    print("case: mlp forward rows", rows, flush=True)
    m = E.Mlp(DIMS, max_batch=256)
    p = m.init_params(seed=rows)
    x, _ = _inputs(rows, rows)
    y = m.forward(x)  # rows > max_batch runs in chunks
    yo, _ = M.forward(p, DIMS, x, emulate_bf16=True)
    assert rel_l2(y, yo) <= 2e-3
    yf, _ = M.forward(p, DIMS, x, emulate_bf16=False)
    assert rel_l2(y, yf) <= 3e-2
    m.close()


@pytest.mark.parametrize("rows", [64, 200])
def test_gradients_match_fp64_oracle(rows):
    print("case: mlp gradients rows", rows, flush=True)
    m = E.Mlp(DIMS, max_batch=256)
    p = m.init_params(seed=7)
    # non-zero biases so every term of the bias path is exercised
    rng = np.random.default_rng(11)
    p = p + (rng.normal(0, 0.05, p.size) * (np.abs(p) == 0)).astype(np.float32)
    m.set_params(p)
    x, t = _inputs(rows, 5)
    loss, g = m.gradients(x, t)
    lo, go = M.loss_and_gradients(p, DIMS, x, t, emulate_bf16=True)
    lf, gf = M.loss_and_gradients(p, DIMS, x, t, emulate_bf16=False)
    assert loss == pytest.approx(lo, rel=2e-3)
    o = 0
    for i, n in zip(DIMS[:-1], DIMS[1:]):
        for size in (n * i, n):
            seg = slice(o, o + size)
            assert rel_l2(g[seg], go[seg]) <= 1e-2, (i, n, size, rel_l2(g[seg], go[seg]))
            assert rel_l2(g[seg], gf[seg]) <= 6e-2, (i, n, size, rel_l2(g[seg], gf[seg]))
            o += size
    assert o == g.size
    m.close()


def test_params_round_trip_and_adam_step():
    print("case: mlp params and one Adam step", flush=True)
    m = E.Mlp(DIMS, max_batch=128)
    p = m.init_params(seed=1)
    assert np.array_equal(m.params(), p)
    x, t = _inputs(128, 2)
    _, g = m.gradients(x, t)
    lr = 1e-3
    m.set_adam(0.9, 0.999, 1e-8, 0.0)
    m.train_step(x, t, lr)
    # first Adam step: m_hat = g, v_hat = g^2, so every parameter with g != 0 moves by ~lr sign(g)
    step = m.params() - p
    moving = np.abs(g) > 1e-6
    assert np.allclose(step[moving], -lr * np.sign(g[moving]), rtol=1e-2, atol=1e-6)
    m.close()


def test_training_reduces_loss():
    print("case: mlp training", flush=True)
    rng = np.random.default_rng(4)
    n = 4096
    x = rng.uniform(0.0, 1.5, (n, DIMS[0])).astype(np.float32)
    A = rng.normal(0, 1.0, (DIMS[0], DIMS[-1]))
    t = np.tanh(x @ A / 3.0).astype(np.float32)
    m = E.Mlp(DIMS, max_batch=512)
    m.init_params(seed=3)
    m.set_adam(0.9, 0.999, 1e-8, 0.0)
    m.set_dataset(x, t)
    losses = m.train(batch=256, epochs=40, lr_start=2e-3, lr_end=1e-4, seed=9)
    assert losses[-1] < losses[0] / 10.0, losses
    y = m.forward(x[:512])
    assert np.mean((y - t[:512]) ** 2) < losses[0] / 10.0
    ms = m.time_forward(512, 5)
    assert ms > 0
    m.close()
