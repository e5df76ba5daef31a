"""TEST INFRASTRUCTURE ONLY (part of the oracle): the per-node MLP of
include/vef_mlp.h restated in numpy — the same layers, the same parameter
layout, the same points where values are rounded to BF16 (every matrix-product
operand), FP32 elsewhere — with its loss and back-propagated gradient.

`mlp(..., bf16=False, dtype=np.float64)` is the exact-arithmetic version the
tests check against central differences, which pins the back-propagation
formulas; the BF16 version is the GPU's checker."""
from __future__ import annotations

import numpy as np

K_IN, WIDTH, K_OUT = 32, 128, 16
SIZES = [("W1", (K_IN, WIDTH)), ("W2", (WIDTH, WIDTH)), ("b2", (WIDTH,)), ("W3", (WIDTH, WIDTH)), ("b3", (WIDTH,)),
         ("W4", (WIDTH, K_OUT)), ("b4", (K_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)
GELU_C = 0.7978845608028654


def unpack(P):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(P[o:o + n]).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name]).reshape(-1) for name, _ in SIZES])


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GELU_C * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * 0.044715 * z * z)


def _inputs(X, dtype):
    n, n_in = X.shape
    A0 = np.zeros((n, K_IN), dtype=dtype)
    A0[:, :n_in] = X
    A0[:, K_IN - 1] = 1.0
    return A0


def forward(P, X, n_out, bf16_ops=True, dtype=np.float32):
    """Y[n][n_out] and the intermediate values (for backward)."""
    p = {k: v.astype(dtype) for k, v in unpack(P).items()}
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, dtype=dtype))
    A = [r(_inputs(np.asarray(X, dtype=dtype), dtype))]
    Z = [None]
    for l, (W, b) in enumerate([("W1", None), ("W2", "b2"), ("W3", "b3")], start=1):
        z = A[-1] @ r(p[W])
        if b:
            z = z + p[b]
        Z.append(z.astype(dtype))
        A.append(r(gelu(z).astype(dtype)))
    Y = (A[-1] @ r(p["W4"]) + p["b4"]).astype(dtype)
    return Y[:, :n_out], (A, Z, Y)


def loss_and_grad(P, X, T, n_out, bf16_ops=True, dtype=np.float32):
    """mean((Y - T)^2) and its gradient in the parameter layout."""
    p = {k: v.astype(dtype) for k, v in unpack(P).items()}
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, dtype=dtype))
    Yn, (A, Z, Y) = forward(P, X, n_out, bf16_ops, dtype)
    n = X.shape[0]
    diff = Yn - np.asarray(T, dtype=dtype)
    loss = float(np.sum(diff.astype(np.float64) ** 2) / (n * n_out))
    dY = np.zeros((n, K_OUT), dtype=dtype)
    dY[:, :n_out] = (2.0 / (n * n_out)) * diff
    dZ = {4: r(dY)}
    for l, W in ((4, "W4"), (3, "W3"), (2, "W2")):
        dA = dZ[l] @ r(p[W]).T
        dZ[l - 1] = r((dA * gelu_grad(Z[l - 1])).astype(dtype))
    g = {}
    for l, (W, b) in enumerate([("W1", None), ("W2", "b2"), ("W3", "b3"), ("W4", "b4")], start=1):
        g[W] = (A[l - 1].T @ dZ[l]).astype(dtype)
        if b:
            g[b] = dZ[l].sum(axis=0).astype(dtype)
    return loss, pack(g)


def init_params(n_in, n_out, seed):
    """N(0, 1/fan_in) weights, zero biases (a numpy draw; the GPU's own
    initialisation uses its own generator)."""
    rng = np.random.default_rng(seed)
    d = {name: np.zeros(shape, dtype=np.float32) for name, shape in SIZES}
    d["W1"][:n_in] = rng.standard_normal((n_in, WIDTH)) / np.sqrt(n_in)
    d["W2"][:] = rng.standard_normal((WIDTH, WIDTH)) / np.sqrt(WIDTH)
    d["W3"][:] = rng.standard_normal((WIDTH, WIDTH)) / np.sqrt(WIDTH)
    d["W4"][:, :n_out] = rng.standard_normal((WIDTH, n_out)) / np.sqrt(WIDTH)
    return pack(d).astype(np.float32)
