"""Pins the CPU oracle of the learned preconditioner: its backward pass against central
finite differences of its own loss, the adjoint identities of the graph operators, the
bfloat16 rounding, and the zero-decoder start (z = block Jacobi)."""
import numpy as np
import pytest

import gnn_oracle as G


def tiny_graph(n_side=(7, 6, 5), seed=3):
    rng = np.random.default_rng(seed)
    nx, ny, nz = n_side
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1)
    X = g.reshape(-1, 3).astype(np.float64) + 0.05 * rng.uniform(-1, 1, (nx * ny * nz, 3))
    n = X.shape[0]
    delta = 2.2
    d = np.linalg.norm(X[:, None, :] - X[None, :, :], axis=2)
    adj = (d <= delta) & ~np.eye(n, dtype=bool)
    cnt = adj.sum(1)
    row_ptr = np.zeros(n + 1, np.int64)
    np.cumsum(cnt, out=row_ptr[1:])
    nbr = np.nonzero(adj)[1]
    L = d[adj]
    near_ptr, near = G.near_graph(row_ptr, nbr, L, delta)
    cmask = np.zeros((n, 3), bool)
    cmask[X[:, 1] < 0.5] = True
    A = rng.standard_normal((n, 3, 3))
    dinv = (A + A.transpose(0, 2, 1) + 6 * np.eye(3)).reshape(n, 9)
    stat = G.static_features(X, row_ptr, near_ptr)
    return dict(near_ptr=near_ptr, near=near, dinv=dinv, cmask=cmask, stat=stat), n


def samples(n, k, seed):
    rng = np.random.default_rng(seed)
    return [(rng.standard_normal((n, 3)), rng.standard_normal((n, 3))) for _ in range(k)]


def test_near_graph_is_symmetric():
    g, n = tiny_graph()
    pairs = set()
    for i in range(n):
        for j in g["near"][g["near_ptr"][i]:g["near_ptr"][i + 1]]:
            pairs.add((i, int(j)))
    assert all((j, i) in pairs for i, j in pairs)
    assert len(pairs) > 0


def test_adjoints():
    g, n = tiny_graph()
    rng = np.random.default_rng(0)
    H, dA = rng.standard_normal((n, 5)), rng.standard_normal((n, 5))
    lhs = np.sum(G.aggregate(H, g["near_ptr"], g["near"]) * dA)
    rhs = np.sum(H * G.aggregate_T(dA, g["near_ptr"], g["near"]))
    assert lhs == pytest.approx(rhs, rel=1e-12)
    for blk in (G.BLK1, 8):
        lhs = np.sum(G.block_mean(H, blk) * dA)
        rhs = np.sum(H * G.block_mean_T(dA, blk))
        assert lhs == pytest.approx(rhs, rel=1e-12)


def test_gradient_matches_finite_differences():
    g, n = tiny_graph()
    theta = G.init_params(7)
    rng = np.random.default_rng(1)
    # a non-zero decoder so every layer receives gradient
    p = G.unflatten(theta)
    p["W_d"] = rng.standard_normal(p["W_d"].shape) * 0.1
    theta = G.flatten(p)
    S = samples(n, 2, 5)
    c = 3.0
    loss, grad = G.loss_and_grad(theta, S, g, c)
    h = 1e-6
    # directional derivatives along random unit directions (a short step crosses no relu
    # kink)
    for k in range(4):
        d = rng.standard_normal(theta.shape)
        d /= np.linalg.norm(d)
        lp, _ = G.loss_and_grad(theta + h * d, S, g, c, want_grad=False)
        lm, _ = G.loss_and_grad(theta - h * d, S, g, c, want_grad=False)
        fd = (lp - lm) / (2 * h)
        assert fd == pytest.approx(np.dot(grad, d), rel=1e-6, abs=1e-9)
    # single coordinates of every parameter block
    o = 0
    for name, shape in G.param_layout():
        k = int(np.prod(shape))
        idx = o + rng.integers(k)
        e = np.zeros_like(theta)
        e[idx] = 1.0
        lp, _ = G.loss_and_grad(theta + h * e, S, g, c, want_grad=False)
        lm, _ = G.loss_and_grad(theta - h * e, S, g, c, want_grad=False)
        assert (lp - lm) / (2 * h) == pytest.approx(grad[idx], rel=1e-4, abs=1e-8), name
        o += k


def test_zero_decoder_is_block_jacobi():
    g, n = tiny_graph()
    theta = G.init_params(0)
    y = np.random.default_rng(2).standard_normal((n, 3))
    z = G.apply(theta, y, g, c=10.0)
    q = G.apply_dinv(g["dinv"], y)
    np.testing.assert_allclose(z[~g["cmask"]], q[~g["cmask"]], rtol=0, atol=0)
    np.testing.assert_array_equal(z[g["cmask"]], y[g["cmask"]])


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-20])
    r = G.bf16(vals)
    # 8 significant bits: exact for 1 + k/128; ties to even
    assert r[0] == 1.0 and r[1] == 1.0  # 1 + 2^-8 is a tie -> even (1.0)
    assert r[2] == 1.0078125
    assert abs(r[3] - vals[3]) <= abs(vals[3]) * 2 ** -8
    assert np.all(np.abs(r - vals) <= np.abs(vals) * 2 ** -8)
