"""CPU oracle of the learned Newton-Krylov preconditioner (test infrastructure only).

The preconditioner has no counterpart in Peridynamics.jl: the reference solves J du = -r
with IterativeSolvers.gmres! and no preconditioner (src/time_solvers/newton_krylov.jl:
448-450).  This module restates, in float64 numpy, the network that libpdx evaluates and
trains on the GPU, so that the device forward pass, loss and gradients can be checked.
The gradients of this module are themselves pinned by central finite differences
(tests/test_gnn_oracle.py).

Network (per point i of the body, in the library's internal Morton order):
  inputs   F_i (16): D^-1 y / s (3), y / rms(y) (3), constrained mask (3),
           normalised reference position (3), near-degree ratio (1), bond-count ratio (1),
           two zeros; s = rms of D^-1 y over free dofs
  encoder  H_i = relu(W_e^T F_i + b_e)                                   (16 -> 64)
  3 layers A_i = mean_{j in N(i)} H_j          (near bonds: L <= 0.6 delta)
           P_i, Q_i, G = mean of H over the 64-point block, the 4096-point block of i
                         (consecutive internal indices) and the whole body
           Z_i = W_l^T [H_i, A_i, P_i, Q_i, G] + b_l                      (320 -> 64)
           H_i <- H_i + relu(Z_i)
  decoder  O_i = W_d^T H_i + b_d                                          (64 -> 3)
  output   z = D^-1 y + c s O on free dofs, z = y on constrained dofs

Training pair (y, x) with J x = y: target t = (x - D^-1 y) / (c s) on free dofs, loss
sum_free (O - t)^2 / sum_free t^2, averaged over the pairs of a batch.

The device stores H, the concatenated layer inputs and the three layer weight matrices in
bfloat16 and accumulates in float32 (encoder and decoder weights stay float32);
`bf16_mode=True` reproduces those roundings here.
"""
import numpy as np

F_IN = 16
F = 64
K_CAT = 5 * F
N_LAYERS = 3
BLK1 = 64
BLK2 = 4096


def param_layout():
    """(name, shape) in the order of the flat float32 parameter vector."""
    lay = [("W_e", (F_IN, F)), ("b_e", (F,))]
    for l in range(1, N_LAYERS + 1):
        lay += [(f"W_{l}", (K_CAT, F)), (f"b_{l}", (F,))]
    lay += [("W_d", (F, 3)), ("b_d", (3,))]
    return lay


def n_params():
    return sum(int(np.prod(s)) for _, s in param_layout())


def unflatten(theta):
    out, o = {}, 0
    for name, shape in param_layout():
        k = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + k], np.float64).reshape(shape)
        o += k
    return out


def flatten(p):
    return np.concatenate([np.asarray(p[name], np.float64).reshape(-1)
                           for name, _ in param_layout()])


def init_params(seed):
    """He-style initialisation; the decoder starts at zero (z = D^-1 y, block Jacobi)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in param_layout():
        if name.startswith("W_") and name != "W_d":
            p[name] = rng.standard_normal(shape) * np.sqrt(2.0 / shape[0]) * (
                0.5 if name != "W_e" else 1.0)
        else:
            p[name] = np.zeros(shape)
    return flatten(p)


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a32 = np.asarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _rd(a, on):
    return bf16(a) if on else a


def block_mean(H, blk):
    n = H.shape[0]
    nb = (n + blk - 1) // blk
    ids = np.arange(n) // blk
    s = np.zeros((nb, H.shape[1]))
    np.add.at(s, ids, H)
    cnt = np.bincount(ids, minlength=nb).astype(np.float64)
    return (s / cnt[:, None])[ids]


def block_mean_T(dP, blk):
    """Adjoint of block_mean: dH_k = (1/|B|) sum_{i in B(k)} dP_i."""
    return block_mean(dP, blk)


def aggregate(H, nbr_ptr, nbr):
    n = H.shape[0]
    deg = np.diff(nbr_ptr)
    A = np.zeros_like(H)
    owner = np.repeat(np.arange(n), deg)
    np.add.at(A, owner, H[nbr])
    with np.errstate(invalid="ignore", divide="ignore"):
        A = np.where(deg[:, None] > 0, A / np.maximum(deg, 1)[:, None], 0.0)
    return A


def aggregate_T(dA, nbr_ptr, nbr):
    """Adjoint of aggregate: dH_j = sum_{i: j in N(i)} dA_i / deg_i."""
    n = dA.shape[0]
    deg = np.diff(nbr_ptr)
    owner = np.repeat(np.arange(n), deg)
    w = dA / np.maximum(deg, 1)[:, None]
    dH = np.zeros_like(dA)
    np.add.at(dH, nbr, w[owner])
    return dH


def near_graph(row_ptr, nbr, L, delta, frac=0.6):
    """Near-bond CSR (bonds with L <= frac * delta), in bond order."""
    keep = L <= frac * delta
    n = row_ptr.shape[0] - 1
    owner = np.repeat(np.arange(n), np.diff(row_ptr))
    cnt = np.bincount(owner[keep], minlength=n)
    ptr = np.zeros(n + 1, np.int64)
    np.cumsum(cnt, out=ptr[1:])
    return ptr, nbr[keep].astype(np.int64)


def static_features(X, row_ptr, near_ptr):
    """Per point: normalised position (3), near-degree ratio, bond-count ratio."""
    lo, hi = X.min(0), X.max(0)
    half = np.maximum(0.5 * (hi - lo), 1e-300)
    Xh = (X - 0.5 * (hi + lo)) / half
    deg = np.diff(near_ptr).astype(np.float64)
    cnt = np.diff(row_ptr).astype(np.float64)
    return np.column_stack([Xh, deg / max(deg.max(), 1.0), cnt / max(cnt.max(), 1.0)])


def apply_dinv(dinv, y):
    """Block Jacobi: q_i = R_i y_i, R_i stored column-major 3x3 (9 per point)."""
    R = dinv.reshape(-1, 3, 3)  # R[i, col, row]
    return np.einsum("icr,ic->ir", R, y.reshape(-1, 3))


def input_features(y, q, cmask, stat):
    free = ~cmask
    s = np.sqrt(np.sum(q[free] ** 2) / max(free.sum(), 1)) + 1e-300
    ry = np.sqrt(np.sum(y[free] ** 2) / max(free.sum(), 1)) + 1e-300
    n = y.shape[0]
    Fi = np.zeros((n, F_IN))
    Fi[:, 0:3] = np.where(free, q / s, 0.0)
    Fi[:, 3:6] = np.where(free, y / ry, 0.0)
    Fi[:, 6:9] = cmask
    Fi[:, 9:14] = stat
    return Fi, s


def forward(theta, Fi, near_ptr, near, bf16_mode=False, keep=False):
    p = unflatten(theta)
    rd = lambda a: _rd(a, bf16_mode)
    We, be = p["W_e"], p["b_e"]  # the encoder and decoder run in float32 on the device
    Z0 = Fi @ We + be
    H = rd(np.maximum(Z0, 0.0))
    cache = {"F": Fi, "Z0": Z0, "X": [], "Z": [], "H": [H]}
    for l in range(1, N_LAYERS + 1):
        A = aggregate(H, near_ptr, near)
        P = block_mean(H, BLK1)
        Q = block_mean(H, BLK2)
        G = np.broadcast_to(H.mean(0), H.shape)
        X = rd(np.concatenate([H, A, P, Q, G], axis=1))
        Z = X @ rd(p[f"W_{l}"]) + p[f"b_{l}"]
        H = rd(H + np.maximum(Z, 0.0))
        if keep:
            cache["X"].append(X)
            cache["Z"].append(Z)
            cache["H"].append(H)
    O = H @ p["W_d"] + p["b_d"]
    return (O, cache) if keep else O


def backward(theta, cache, dO, near_ptr, near, bf16_mode=False):
    """Gradient of sum(O * dO) with respect to theta.  In bf16 mode the roundings are
    passed straight through and the layer weights enter dX as their bfloat16 values, as on
    the device."""
    p = unflatten(theta)
    g = {}
    H = cache["H"][-1]
    g["W_d"] = H.T @ dO
    g["b_d"] = dO.sum(0)
    dH = dO @ p["W_d"].T
    n = H.shape[0]
    for l in range(N_LAYERS, 0, -1):
        X, Z = cache["X"][l - 1], cache["Z"][l - 1]
        dZ = dH * (Z > 0)
        g[f"W_{l}"] = X.T @ dZ
        g[f"b_{l}"] = dZ.sum(0)
        dX = dZ @ _rd(p[f"W_{l}"], bf16_mode).T
        dHin = dH + dX[:, 0:F]
        dHin += aggregate_T(dX[:, F:2 * F], near_ptr, near)
        dHin += block_mean_T(dX[:, 2 * F:3 * F], BLK1)
        dHin += block_mean_T(dX[:, 3 * F:4 * F], BLK2)
        dHin += np.broadcast_to(dX[:, 4 * F:5 * F].sum(0) / n, dHin.shape)
        dH = dHin
    dZ0 = dH * (cache["Z0"] > 0)
    g["W_e"] = cache["F"].T @ dZ0
    g["b_e"] = dZ0.sum(0)
    return flatten(g)


def sample_target(y, x, q, cmask, s, c):
    free = ~cmask
    return np.where(free, (x - q) / (c * s), 0.0)


def loss_and_grad(theta, samples, graph, c, bf16_mode=False, want_grad=True):
    """samples: list of (y, x) with y, x of shape (n, 3); graph: dict with near_ptr, near,
    dinv, cmask (n, 3 bool), stat (n, 5).  Returns (loss, grad)."""
    near_ptr, near = graph["near_ptr"], graph["near"]
    total = 0.0
    grad = np.zeros(n_params())
    for y, x in samples:
        q = apply_dinv(graph["dinv"], y)
        Fi, s = input_features(y, q, graph["cmask"], graph["stat"])
        t = sample_target(y, x, q, graph["cmask"], s, c)
        O, cache = forward(theta, Fi, near_ptr, near, bf16_mode, keep=True)
        free = ~graph["cmask"]
        den = max(np.sum(t[free] ** 2), 1e-300)
        r = np.where(free, O - t, 0.0)
        total += np.sum(r ** 2) / den
        if want_grad:
            grad += backward(theta, cache, 2.0 * r / den, near_ptr, near, bf16_mode)
    B = len(samples)
    return total / B, grad / B


def apply(theta, y, graph, c, bf16_mode=False):
    """z = M(y) (the preconditioner)."""
    q = apply_dinv(graph["dinv"], y)
    Fi, s = input_features(y, q, graph["cmask"], graph["stat"])
    O = forward(theta, Fi, graph["near_ptr"], graph["near"], bf16_mode)
    return np.where(graph["cmask"], y, q + c * s * O)
