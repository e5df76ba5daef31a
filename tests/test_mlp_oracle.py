"""The MLP restatement (oracle/mlp_ref.py) pinned by central differences: in
float64 without BF16 rounding, its back-propagated gradient of mean((Y-T)^2)
equals the numerical derivative for weights of every layer, the bias row of
the first layer and the later biases; a planted slip in the GELU derivative
fails the same check."""
import numpy as np
import pytest

import mlp_ref


@pytest.fixture(scope="module")
def problem():
    rng = np.random.default_rng(7)
    n, n_in, n_out = 9, 16, 6
    P = mlp_ref.init_params(n_in, n_out, 3).astype(np.float64)
    P += 0.05 * rng.standard_normal(P.size) * (P != 0)
    d = mlp_ref.unpack(P)
    for b in ("b2", "b3", "b4"):
        d[b][:] = 0.1 * rng.standard_normal(d[b].shape)
    d["b4"][n_out:] = 0.0
    d["W1"][31] = 0.1 * rng.standard_normal(128)
    P = mlp_ref.pack(d)
    X = rng.standard_normal((n, n_in))
    T = rng.standard_normal((n, n_out))
    return P, X, T, n_out


def _fd_check(P, X, T, n_out, grad_fn):
    loss, g = grad_fn(P)
    rng = np.random.default_rng(11)
    d = mlp_ref.unpack(np.arange(P.size))
    idx = np.concatenate([rng.choice(d["W1"][:16].ravel(), 8), d["W1"][31][:4], rng.choice(d["W2"].ravel(), 8),
                          d["b2"][:3], rng.choice(d["W3"].ravel(), 8), d["b3"][:3],
                          rng.choice(d["W4"][:, :n_out].ravel(), 8), d["b4"][:n_out]])
    worst = 0.0
    for i in idx:
        h = 1e-6
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (mlp_ref.loss_and_grad(Pp, X, T, n_out, False, np.float64)[0]
              - mlp_ref.loss_and_grad(Pm, X, T, n_out, False, np.float64)[0]) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / (abs(fd) + 1e-8))
    return worst


def test_gradient_matches_central_differences(problem):
    P, X, T, n_out = problem
    worst = _fd_check(P, X, T, n_out, lambda p: mlp_ref.loss_and_grad(p, X, T, n_out, False, np.float64))
    assert worst < 1e-5


def test_planted_slip_in_gelu_derivative_fails(problem, monkeypatch):
    P, X, T, n_out = problem
    monkeypatch.setattr(mlp_ref, "gelu_grad", lambda z: 0.5 * (1.0 + np.tanh(mlp_ref.GELU_C * z)))
    worst = _fd_check(P, X, T, n_out, lambda p: mlp_ref.loss_and_grad(p, X, T, n_out, False, np.float64))
    assert worst > 1e-3


def test_bf16_rounding():
    a = np.array([1.0, 1.00390625, 1.005859375, -2.0e-3, 3.0e38], dtype=np.float32)
    r = mlp_ref.bf16(a)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125  # ties to even, then up
    assert np.all(np.abs(r - a) <= np.abs(a) * 2.0 ** -8)
