"""Learned preconditioner on the device against its CPU oracle (oracle/gnn_oracle.py): the
input features, the network's coarse fields, the least-squares loss and its gradients at
the same parameters and samples, and the gain of the trained network's coarse space over
the two-level preconditioner of the translations alone.

Tolerances: the device keeps activations and layer weights in bfloat16 (8 significant
bits) and sums in float32; the oracle reproduces the roundings (bf16_mode) but not the
summation order, so a value that lands next to a bfloat16 rounding boundary can round the
other way.  Features within 1e-6; fields within 2e-2 normwise; loss within 1e-3; every
gradient block within 3e-2 normwise (cosine > 0.999)."""
import sys

import numpy as np
import pytest

import peridynamics as pd
import gnn_oracle as G

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def block(nx, ny, nz, dx=1e-3, top=1e-4, seed=11):
    pos, vol = pd.lattice_block(nx, ny, nz, dx, jitter_amplitude=0.05, seed=seed)
    b = pd.Body(pd.BBMaterial(), pos, vol)
    pd.material(b, horizon=3.015 * dx, rho=2700.0, E=70e9)
    y = pos[1]
    pd.point_set(b, "bottom", y < y.min() + 0.5 * dx)
    pd.point_set(b, "top", y > y.max() - 0.5 * dx)
    for d in "xyz":
        pd.displacement_bc(lambda p: 0.0, b, "bottom", d)
    pd.displacement_bc(lambda p: top, b, "top", "y")
    return b


def device_graph(dh):
    perm, near_ptr, near, feat, dinv, cmask = dh.ctx.gnn_graph()
    agg, nagg, _ = dh.ctx.coarse_info()
    assert agg >= G.agg_size(perm.size) and nagg == (perm.size + agg - 1) // agg
    return dict(near_ptr=near_ptr, near=near.astype(np.int64), cmask=cmask,
                feat=feat.astype(np.float64), agg=agg), perm


def random_theta(seed):
    theta = G.init_params(seed)
    p = G.unflatten(theta)
    rng = np.random.default_rng(seed + 1)
    p["b_d"] = rng.standard_normal(G.N_OUT) * 0.05
    for l in range(1, G.N_LAYERS + 1):
        p[f"b_{l}"] = rng.standard_normal(G.F) * 0.05
    return G.flatten(p).astype(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def small():
    b = block(20, 20, 6)
    dh = pd.MI355XBodyDataHandler(b)
    dh.init_newton_krylov(pd.NewtonKrylov(steps=1))
    g, perm = device_graph(dh)
    yield b, dh, g, perm
    dh.close()


def test_graph_and_features_match(need_gpu, small):
    say("gnn graph and features")
    b, dh, g, perm = small
    row_ptr, nbr, L, _, _ = dh.bonds()
    iperm = np.empty_like(perm)
    iperm[perm] = np.arange(perm.size)
    # near bonds of internal point m = bonds of caller point perm[m] with L <= 0.6 delta
    delta = b.point_params[0].delta
    for m in range(0, b.n_points, 97):
        i = perm[m]
        sel = L[row_ptr[i]:row_ptr[i + 1]] <= 0.6 * delta
        want = np.sort(iperm[nbr[row_ptr[i]:row_ptr[i + 1]][sel]])
        got = np.sort(g["near"][g["near_ptr"][m]:g["near_ptr"][m + 1]])
        np.testing.assert_array_equal(got, want)
    # the 16 input features from the reference positions in internal order
    cnt = np.diff(row_ptr)[perm]
    rp = np.zeros(b.n_points + 1, np.int64)
    np.cumsum(cnt, out=rp[1:])
    want = G.input_features(b.xyz[perm], rp, g["near_ptr"], g["cmask"], g["agg"])
    np.testing.assert_allclose(g["feat"], want, rtol=0, atol=1e-6)


def test_fields_match_oracle(need_gpu, small):
    say("gnn fields")
    b, dh, g, perm = small
    theta = random_theta(3)
    dh.ctx.gnn_set_params(theta, 1.0)
    f = dh.ctx.gnn_fields().astype(np.float64)
    fo = G.fields(theta, g, bf16_mode=True)
    err = np.linalg.norm(f - fo) / np.linalg.norm(fo)
    say(f"  fields relative error {err:.2e}")
    assert err < 2e-2


def test_gradients_match_oracle(need_gpu, small):
    say("gnn loss and gradients")
    b, dh, g, perm = small
    theta = random_theta(5)
    dh.ctx.gnn_set_params(theta, 1.0)
    E = np.random.default_rng(6).standard_normal((3, b.n_points, 3))
    E = np.where(g["cmask"][None], 0.0, E)
    loss, grad = dh.ctx.gnn_loss_grad(E)
    lo, go = G.loss_and_grad(theta, E, g, bf16_mode=True)
    say(f"  loss device {loss:.6e} oracle {lo:.6e}")
    assert loss == pytest.approx(lo, rel=1e-3)
    o = 0
    for name, shape in G.param_layout():
        k = int(np.prod(shape))
        gd, gr = grad[o:o + k].astype(np.float64), go[o:o + k]
        rel = np.linalg.norm(gd - gr) / max(np.linalg.norm(gr), 1e-30)
        cos = gd @ gr / max(np.linalg.norm(gd) * np.linalg.norm(gr), 1e-30)
        say(f"  {name}: relative error {rel:.2e}, cosine {cos:.6f}")
        assert rel < 3e-2 and cos > 0.999, name
        o += k


def test_training_lowers_loss(need_gpu):
    """Adam steps on the error samples of one load step solved by the translations'
    two-level preconditioner lower the loss, and the held-out loss with the fields is
    below that of the translations alone."""
    say("gnn training loss")
    b = block(16, 16, 8)
    dh = pd.MI355XBodyDataHandler(b)
    nk = pd.NewtonKrylov(steps=4, tol=1e-30, rel_tol=1e-8, maxiter=20, gmres_restart=30,
                         gmres_maxiter=300, precond="two_level")
    dh.init_newton_krylov(nk)
    dh.ctx.gnn_collect(True, 256)
    dh.nk_load_step(nk, 1)
    dh.ctx.gnn_collect(False)
    nn, ns = dh.ctx.gnn_pairs()
    say(f"  error samples {ns}")
    assert ns >= 20
    l0 = dh.ctx.gnn_train(1, batch=64, lr=1e-12, seed=1)
    l1 = dh.ctx.gnn_train(150, batch=64, lr=3e-3, seed=2)
    ratio, use = dh.ctx.gnn_gate()
    say(f"  loss before {l0:.4f} after {l1:.4f}; held-out ratio fields / translations {ratio:.3f}")
    assert l1 < 0.9 * l0
    assert ratio < 1.0
    dh.close()


def test_learned_preconditioner_gain(need_gpu):
    """Load steps 1-3 are solved with the two-level preconditioner of the translations while
    their error samples are collected; the network is trained on them (the default 40 Adam
    steps) and its coarse space preconditions load steps 4-10.  The learned preconditioner needs fewer Krylov iterations
    than the two-level one without the network (strict gain) and than block Jacobi, uses no
    fallback, and every solve reaches the 1e-8 relative Newton tolerance."""
    say("gnn gain")
    res = {}
    for pre in ("jacobi", "two_level", "learned"):
        b = block(30, 30, 8)
        dh = pd.MI355XBodyDataHandler(b)
        nk = pd.NewtonKrylov(steps=10, tol=1e-30, rel_tol=1e-8, maxiter=30, gmres_restart=30,
                             gmres_maxiter=300, precond=pre, learn_steps=3, krylov_pairs=512)
        stats = dh.solve_nk(nk)
        late = stats[3:]
        res[pre] = (sum(s.krylov_iters for s in late), sum(s.seconds for s in late),
                    sum(s.precond_fallbacks for s in late), dh.get("displacement"))
        for s in late:
            assert s.residual_final < 1e-8 * s.residual_first
        say(f"  {pre}: krylov {res[pre][0]} in {res[pre][1]:.3f} s, fallbacks {res[pre][2]}"
            + (f", training {dh.train_info}" if getattr(dh, "train_info", None) else ""))
        dh.close()
    say(f"  network gain over two-level: {res['two_level'][0] / max(res['learned'][0], 1):.2f}x "
        f"iterations; learned over Jacobi: {res['jacobi'][0] / max(res['learned'][0], 1):.2f}x")
    assert res["learned"][2] == 0
    assert res["learned"][0] < res["two_level"][0]
    assert res["learned"][0] < res["jacobi"][0]
    uj = res["jacobi"][3]
    for pre in ("two_level", "learned"):
        assert np.linalg.norm(res[pre][3] - uj) / np.linalg.norm(uj) < 1e-6
