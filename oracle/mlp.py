"""CPU FP64 restatement of the surrogate MLP: test infrastructure only (the checker of the HIP
forward / backward kernels, never part of the product path).

The network is what libemfdtd.so's emfd_mlp_* computes: hidden layers z = W x + b,
h = gelu(z) with the tanh form 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))), a linear output
layer, and the loss mean((y - t)^2) over batch x n_out. Parameters are a flat vector: per
layer W [n_out][n_in] (row-major), then b [n_out].

`emulate_bf16=True` rounds to bfloat16 (round to nearest even) exactly where the device does:
inputs and weights before every product, hidden activations after GELU, and the loss and
pre-activation gradients before they enter a product (their FP32 values feed the bias sums).
Everything else is FP64. With `emulate_bf16=False` this is the plain FP64 network, whose
gradient the finite-difference test pins.

No reference counterpart exists (the reference has no learned part).
"""
from __future__ import annotations

import numpy as np

SQRT_2_OVER_PI = 0.7978845608028654


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even) and return as float64."""
    f = np.asarray(a, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(SQRT_2_OVER_PI * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(SQRT_2_OVER_PI * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * SQRT_2_OVER_PI * (1.0 + 3.0 * 0.044715 * z * z)


def unpack(params, dims):
    out, o = [], 0
    p = np.asarray(params, np.float64)
    for i, n in zip(dims[:-1], dims[1:]):
        W = p[o:o + n * i].reshape(n, i)
        o += n * i
        b = p[o:o + n]
        o += n
        out.append((W, b))
    assert o == p.size
    return out


def forward(params, dims, x, emulate_bf16=True):
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params, dims)
    a = r(x)
    cache = []
    for li, (W, b) in enumerate(layers):
        z = a @ r(W).T + b
        cache.append((a, z))
        if li < len(layers) - 1:
            a = r(gelu(z))
        else:
            a = z
    return a, cache


def loss_and_gradients(params, dims, x, t, emulate_bf16=True):
    r = bf16 if emulate_bf16 else (lambda v: np.asarray(v, np.float64))
    layers = unpack(params, dims)
    y, cache = forward(params, dims, x, emulate_bf16)
    t = np.asarray(t, np.float64)
    d = y - t
    loss = float(np.mean(d * d))
    G = 2.0 * d / d.size
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        a, _ = cache[li]
        Gb = r(G)
        grads.append((Gb.T @ a, G.sum(axis=0)))
        if li == 0:
            break
        _, zp = cache[li - 1]
        G = (Gb @ r(W)) * gelu_grad(zp)
    flat = []
    for dW, db in reversed(grads):
        flat += [dW.ravel(), db]
    return loss, np.concatenate(flat)


def loss_only(params, dims, x, t, emulate_bf16=False):
    y, _ = forward(params, dims, x, emulate_bf16)
    d = y - np.asarray(t, np.float64)
    return float(np.mean(d * d))
