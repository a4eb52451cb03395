"""Pins of the CPU FP64 MLP restatement (oracle/mlp.py) that checks the HIP surrogate kernels:
its gradient against central finite differences, its bfloat16 rounding against torch's
conversion and known round-to-nearest-even ties, and GELU's derivative against a difference
quotient."""
import numpy as np
import pytest

from oracle import mlp as M


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    dims = (5, 32, 32, 3)
    n = sum(o * i + o for i, o in zip(dims[:-1], dims[1:]))
    p = rng.normal(0, 0.4, n)
    x = rng.normal(0, 1, (7, dims[0]))
    t = rng.normal(0, 1, (7, dims[-1]))
    loss, g = M.loss_and_gradients(p, dims, x, t, emulate_bf16=False)
    assert loss == pytest.approx(M.loss_only(p, dims, x, t), rel=1e-15)
    h = 1e-6
    for k in rng.choice(n, 40, replace=False):
        e = np.zeros(n)
        e[k] = h
        fd = (M.loss_only(p + e, dims, x, t) - M.loss_only(p - e, dims, x, t)) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-9), k


def test_bf16_rounding_known_answers():
    # ties go to the even mantissa: 1 + 2^-8 -> 1, 1 + 3 * 2^-8 -> 1 + 2^-6
    v = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 3.14159265, -2.5e-3, 65504.0])
    r = M.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2.0 ** -6
    assert r[3] == 3.140625
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    a = (rng.standard_normal(20000) * 10.0 ** rng.uniform(-6, 6, 20000)).astype(np.float32)
    want = torch.from_numpy(a).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(M.bf16(a), want)


def test_gelu_grad_is_the_derivative():
    z = np.linspace(-6, 6, 1001)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - M.gelu_grad(z))) < 1e-8
    # GELU(0) = 0, GELU(x) -> x for large x, -> 0 for very negative x
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert M.gelu(np.array([10.0]))[0] == pytest.approx(10.0)
    assert abs(M.gelu(np.array([-10.0]))[0]) < 1e-12
