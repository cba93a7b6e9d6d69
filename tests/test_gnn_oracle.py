"""Pins the CPU oracle of the learned preconditioner: its backward pass against central
finite differences of its own loss, the least-squares loss against closed forms (zero for
errors in the span of the coarse basis, the translations' projection is the per-aggregate
mean), the adjoint identities of the graph operators and the bfloat16 rounding."""
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
    agg = G.agg_size(n)
    feat = G.input_features(X, row_ptr, near_ptr, cmask, agg)
    return dict(near_ptr=near_ptr, near=near, cmask=cmask, feat=feat, agg=agg, X=X), n


def samples(n, k, seed):
    return np.random.default_rng(seed).standard_normal((k, n, 3))


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
    E = samples(n, 3, 5)
    loss, grad = G.loss_and_grad(theta, E, g)
    h = 1e-6
    # directional derivatives along random unit directions (a short step crosses no relu
    # kink)
    for k in range(4):
        d = rng.standard_normal(theta.shape)
        d /= np.linalg.norm(d)
        lp, _ = G.loss_and_grad(theta + h * d, E, g, want_grad=False)
        lm, _ = G.loss_and_grad(theta - h * d, E, g, want_grad=False)
        fd = (lp - lm) / (2 * h)
        assert fd == pytest.approx(np.dot(grad, d), rel=1e-5, abs=1e-9)
    # single coordinates of every parameter block
    o = 0
    for name, shape in G.param_layout():
        k = int(np.prod(shape))
        idx = o + rng.integers(k)
        e = np.zeros_like(theta)
        e[idx] = 1.0
        lp, _ = G.loss_and_grad(theta + h * e, E, g, want_grad=False)
        lm, _ = G.loss_and_grad(theta - h * e, E, g, want_grad=False)
        assert (lp - lm) / (2 * h) == pytest.approx(grad[idx], rel=1e-4, abs=1e-8), name
        o += k


def test_lsq_gradient_matches_finite_differences():
    """d loss / d fields by the envelope theorem, including the ridge's dependence on the
    fields, against central differences in single field entries."""
    g, n = tiny_graph()
    rng = np.random.default_rng(4)
    O = rng.standard_normal((n, G.N_OUT))
    E = samples(n, 2, 9)
    loss, dO = G.lsq_loss(O, E, g["cmask"], g["agg"])
    h = 1e-6
    free_rows = np.where(~g["cmask"].all(1))[0]
    for _ in range(12):
        i, o = int(rng.choice(free_rows)), int(rng.integers(G.N_OUT))
        if g["cmask"][i, o % 3]:
            assert dO[i, o] == 0.0
            continue
        Op, Om = O.copy(), O.copy()
        Op[i, o] += h
        Om[i, o] -= h
        fd = (G.lsq_loss(Op, E, g["cmask"], g["agg"])[0] - G.lsq_loss(Om, E, g["cmask"], g["agg"])[0]) / (2 * h)
        assert fd == pytest.approx(dO[i, o], rel=1e-5, abs=1e-10)


def test_errors_in_the_coarse_span_have_no_loss():
    """Samples that are, on every aggregate, a combination of the translations and the
    fields are reproduced: the loss is at the level of the ridge (1e-6), against ~0.9 for
    random samples."""
    g, n = tiny_graph()
    rng = np.random.default_rng(6)
    O = rng.standard_normal((n, G.N_OUT))
    B = G.basis(O, g["cmask"])
    agg = g["agg"]
    E = np.zeros((3, n, 3))
    for s in range(3):
        for I in range((n + agg - 1) // agg):
            sl = slice(I * agg, min(n, (I + 1) * agg))
            E[s, sl] = np.einsum("pdk,k->pd", B[sl], rng.standard_normal(3 + G.K_F))
    loss, _ = G.lsq_loss(O, E, g["cmask"], agg)
    assert loss < 1e-5
    loss_r, _ = G.lsq_loss(O, samples(n, 3, 2), g["cmask"], agg)
    assert loss_r > 0.5


def test_translations_alone_project_on_the_aggregate_mean():
    """Without fields the fit on an aggregate is the mean of each component over its free
    dofs: the loss is the sum of the squared deviations from those means (to the ridge)."""
    g, n = tiny_graph()
    E = samples(n, 2, 3)
    agg, cm = g["agg"], g["cmask"]
    loss, _ = G.lsq_loss(np.zeros((n, G.N_OUT)), E, cm, agg, use_fields=False)
    want = 0.0
    for e in E:
        e = np.where(cm, 0.0, e)
        tot = 0.0
        for I in range((n + agg - 1) // agg):
            sl = slice(I * agg, min(n, (I + 1) * agg))
            for d in range(3):
                v = e[sl, d][~cm[sl, d]]
                if v.size:
                    tot += np.sum((v - v.mean()) ** 2)
        want += tot / np.sum(e ** 2)
    want /= len(E)
    assert loss == pytest.approx(want, rel=1e-5)


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-20])
    r = G.bf16(vals)
    # 8 significant bits: exact for 1 + k/128; ties to even
    assert r[0] == 1.0 and r[1] == 1.0  # 1 + 2^-8 is a tie -> even (1.0)
    assert r[2] == 1.0078125
    assert abs(r[3] - vals[3]) <= abs(vals[3]) * 2 ** -8
    assert np.all(np.abs(r - vals) <= np.abs(vals) * 2 ** -8)
