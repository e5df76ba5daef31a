"""The per-node MLP on the matrix cores against its numpy restatement
(oracle/mlp_ref.py, the same BF16 rounding points): forward values, the
back-propagated gradient of every parameter block, and training progress."""
import numpy as np
import pytest

import mlp_ref

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def net_and_data():
    from ve_flow_around_cylinder_fortran_amd.learned import NodeMLP

    rng = np.random.default_rng(5)
    n, n_in, n_out = 137, 16, 6  # a partial last tile
    P = mlp_ref.init_params(n_in, n_out, 9)
    d = mlp_ref.unpack(P)
    d["W1"][31] = 0.1 * rng.standard_normal(128).astype(np.float32)
    for b in ("b2", "b3"):
        d[b][:] = 0.1 * rng.standard_normal(128).astype(np.float32)
    d["b4"][:n_out] = 0.1 * rng.standard_normal(n_out).astype(np.float32)
    P = mlp_ref.pack(d).astype(np.float32)
    X = rng.standard_normal((n, n_in)).astype(np.float32)
    T = rng.standard_normal((n, n_out)).astype(np.float32)
    net = NodeMLP(n_in, n_out, seed=1)
    net.set_params(P)
    return net, P, X, T, n_out


def test_parameters_round_trip(net_and_data):
    net, P, X, T, n_out = net_and_data
    assert net.n_params == mlp_ref.N_PARAMS
    assert np.array_equal(net.params(), P)


def test_forward_matches_restatement(net_and_data):
    net, P, X, T, n_out = net_and_data
    print("[gpu mlp] forward", flush=True)
    Y = net.forward(torch.from_numpy(X).cuda())
    torch.cuda.synchronize()
    Yg = Y.cpu().numpy()
    Yc, _ = mlp_ref.forward(P, X, n_out)
    scale = np.abs(Yc).max()
    assert np.abs(Yg - Yc).max() <= 2e-2 * scale
    assert np.abs(Yg - Yc).mean() <= 1e-3 * scale


def test_gradient_matches_restatement(net_and_data):
    net, P, X, T, n_out = net_and_data
    print("[gpu mlp] gradient", flush=True)
    g, loss = net.gradient(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda())
    torch.cuda.synchronize()
    gl, gc = mlp_ref.loss_and_grad(P, X, T, n_out)
    assert abs(loss.item() - gl) <= 1e-3 * gl
    gg = g.cpu().numpy()
    dg, dc = mlp_ref.unpack(gg), mlp_ref.unpack(gc)
    for name in ("W1", "W2", "b2", "W3", "b3", "W4", "b4"):
        err = np.linalg.norm(dg[name] - dc[name]) / (np.linalg.norm(dc[name]) + 1e-30)
        assert err < 2e-2, (name, err)
    # the padding of the first layer's inputs and of the outputs gets no gradient
    assert np.all(dg["W1"][16:31] == 0) and np.all(dg["W4"][:, n_out:] == 0) and np.all(dg["b4"][n_out:] == 0)


def test_adam_reduces_the_loss():
    from ve_flow_around_cylinder_fortran_amd.learned import NodeMLP

    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, (2000, 4)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) * X[:, 1], X[:, 2] ** 2 - X[:, 3]], 1).astype(np.float32)
    net = NodeMLP(4, 2, seed=4)
    Xd, Td = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    first = net.adam_step(Xd, Td, 3e-3).item()
    for _ in range(400):
        last = net.adam_step(Xd, Td, 3e-3)
    torch.cuda.synchronize()
    assert last.item() < 0.05 * first
