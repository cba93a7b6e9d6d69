"""Learned preconditioner on the device against its CPU oracle (oracle/gnn_oracle.py):
forward pass, loss and gradients at the same parameters and inputs, and the gain of the
trained network over the point-block Jacobi preconditioner it falls back to.

Tolerances: the device keeps activations and layer weights in bfloat16 (8 significant
bits) and sums in float32; the oracle reproduces the roundings (bf16_mode) but not the
summation order, so a value that lands next to a bfloat16 rounding boundary can round the
other way.  Forward: network part of z within 2e-2 normwise; loss within 1e-3; every
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
    perm, near_ptr, near, stat, dinv, cmask = dh.ctx.gnn_graph()
    return dict(near_ptr=near_ptr, near=near.astype(np.int64), dinv=dinv, cmask=cmask,
                stat=stat.astype(np.float64)), perm


def random_theta(seed):
    theta = G.init_params(seed)
    p = G.unflatten(theta)
    rng = np.random.default_rng(seed + 1)
    p["W_d"] = rng.standard_normal(p["W_d"].shape) * 0.2
    p["b_d"] = rng.standard_normal(3) * 0.05
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


def test_graph_matches_bond_list(need_gpu, small):
    say("gnn graph")
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


def test_forward_matches_oracle(need_gpu, small):
    say("gnn forward")
    b, dh, g, perm = small
    theta = random_theta(3)
    c = 7.5
    dh.ctx.gnn_set_params(theta, c)
    rng = np.random.default_rng(4)
    for k in range(2):
        y = rng.standard_normal((b.n_points, 3))
        z = dh.ctx.gnn_apply(y)
        zo = G.apply(theta, y, g, c, bf16_mode=True)
        q = G.apply_dinv(g["dinv"], y)
        free = ~g["cmask"]
        err = np.linalg.norm((z - zo)[free]) / np.linalg.norm((zo - q)[free])
        say(f"  forward relative error of the network part {err:.2e}")
        assert err < 2e-2
        np.testing.assert_array_equal(z[g["cmask"]], y[g["cmask"]])


def test_gradients_match_oracle(need_gpu, small):
    say("gnn gradients")
    b, dh, g, perm = small
    theta = random_theta(5)
    c = 4.0
    dh.ctx.gnn_set_params(theta, c)
    rng = np.random.default_rng(6)
    Y = rng.standard_normal((2, b.n_points, 3))
    X = rng.standard_normal((2, b.n_points, 3))
    loss, grad = dh.ctx.gnn_loss_grad(Y, X)
    lo, go = G.loss_and_grad(theta, list(zip(Y, X)), g, c, bf16_mode=True)
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


def test_training_step_lowers_loss(need_gpu):
    """Adam steps on the pairs of one Jacobi-preconditioned load step lower the loss (1.0 for
    the untrained network, whose decoder starts at zero)."""
    say("gnn training loss")
    b = block(16, 16, 6)
    dh = pd.MI355XBodyDataHandler(b)
    nk = pd.NewtonKrylov(steps=4, tol=1e-30, rel_tol=1e-8, maxiter=20, gmres_restart=30,
                         gmres_maxiter=300, precond="jacobi")
    dh.init_newton_krylov(nk)
    dh.ctx.gnn_collect(True, 512)
    dh.nk_load_step(nk, 1)
    dh.ctx.gnn_collect(False)
    nn, nkr = dh.ctx.gnn_pairs()
    say(f"  pairs: newton {nn}, krylov {nkr}")
    assert nn >= 1 and nkr >= 10
    l0 = dh.ctx.gnn_train(1, batch=8, lr=1e-12, seed=1)
    l1 = dh.ctx.gnn_train(300, batch=4, lr=2e-3, seed=2)
    say(f"  loss before {l0:.4f} after {l1:.4f}")
    assert l0 == pytest.approx(1.0, rel=1e-3)
    assert l1 < 0.95 * l0
    dh.close()


def test_learned_preconditioner_gain(need_gpu):
    """Load steps 1-3 are solved with block Jacobi while their Krylov pairs are collected;
    the network is trained on them and preconditions load steps 4-10.  Reported: Krylov
    iterations and time of steps 4-10 for block Jacobi, the two-level preconditioner the
    network builds on, and the learned one.  Every solve reaches the 1e-8 relative Newton
    tolerance; the learned preconditioner needs fewer iterations than block Jacobi and not
    more than 10 % over the two-level one (the network's own gain is the difference)."""
    say("gnn gain")
    res = {}
    for pre in ("jacobi", "two_level", "learned"):
        b = block(30, 30, 6)
        dh = pd.MI355XBodyDataHandler(b)
        nk = pd.NewtonKrylov(steps=10, tol=1e-30, rel_tol=1e-8, maxiter=30, gmres_restart=30,
                             gmres_maxiter=600, precond=pre, learn_steps=3, train_steps=300)
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
    assert res["learned"][0] < res["jacobi"][0]
    assert res["learned"][0] <= 1.1 * res["two_level"][0]
    uj = res["jacobi"][3]
    for pre in ("two_level", "learned"):
        assert np.linalg.norm(res[pre][3] - uj) / np.linalg.norm(uj) < 1e-6
