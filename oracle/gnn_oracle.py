"""CPU oracle of the learned Newton-Krylov preconditioner (test infrastructure only).

The preconditioner has no counterpart in Peridynamics.jl: the reference solves J du = -r
with IterativeSolvers.gmres! and no preconditioner (src/time_solvers/newton_krylov.jl:
448-450).  This module restates, in float64 numpy, the network that libpdx evaluates and
trains on the GPU and the loss it is trained on, so that the device forward pass, loss and
gradients can be checked.  The gradients of this module are themselves pinned by central
finite differences (tests/test_gnn_oracle.py).

Network (per point i of the body, in the library's internal Morton order):
  inputs   F_i (16): normalised reference position (3), position relative to the point's
           aggregate (3), constrained mask (3), near-degree ratio (1), bond-count ratio (1),
           squares of the relative position (3), 1, 0
  encoder  H_i = relu(W_e^T F_i + b_e)                                   (16 -> 64)
  3 layers A_i = mean_{j in N(i)} H_j          (near bonds: L <= 0.6 delta)
           P_i, Q_i, G = mean of H over the 64-point block, the 4096-point block of i
                         (consecutive internal indices) and the whole body
           Z_i = W_l^T [H_i, A_i, P_i, Q_i, G] + b_l                      (320 -> 64)
           H_i <- H_i + relu(Z_i)
  decoder  O_i = W_d^T H_i + b_d                                          (64 -> 24)
  fields   phi_i[f] = O_i[3f : 3f + 3], f < 8

Aggregates are runs of `agg` consecutive internal indices (agg: a power of two >= 64 with at
most 256 aggregates).  The coarse basis of aggregate I is B_I = [e_x, e_y, e_z, phi_0 ..
phi_7] on I's free dofs.  Loss over error samples e_s:
  (1/S) sum_s w_s sum_I min_c ( |e_sI - B_I c|^2 + lambda_I |c|^2 ),
  w_s = 1 / |e_s|^2 (free dofs), lambda_I = 1e-6 tr(B_I^T B_I) / 11 + 1e-30.

The device stores H, the concatenated layer inputs and the three layer weight matrices in
bfloat16 and accumulates in float32 (encoder and decoder weights stay float32);
`bf16_mode=True` reproduces those roundings here.
"""
import numpy as np

F_IN = 16
F = 64
K_F = 8
N_OUT = 3 * K_F
RIDGE = 1e-6
K_CAT = 5 * F
N_LAYERS = 3
BLK1 = 64
BLK2 = 4096


def param_layout():
    """(name, shape) in the order of the flat float32 parameter vector."""
    lay = [("W_e", (F_IN, F)), ("b_e", (F,))]
    for l in range(1, N_LAYERS + 1):
        lay += [(f"W_{l}", (K_CAT, F)), (f"b_{l}", (F,))]
    lay += [("W_d", (F, N_OUT)), ("b_d", (N_OUT,))]
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
    """He-style initialisation; a small random decoder (zero fields get no gradient)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in param_layout():
        if name == "W_d":
            p[name] = rng.standard_normal(shape) * np.sqrt(1.0 / shape[0])
        elif name.startswith("W_"):
            p[name] = rng.standard_normal(shape) * np.sqrt(2.0 / shape[0]) * (
                0.5 if name != "W_e" else 1.0)
        else:
            p[name] = np.zeros(shape)
    return flatten(p)


def agg_size(n):
    """Aggregate size of the two-level preconditioner: a power of two >= 64, at most 256
    aggregates."""
    agg = 64
    while (n + agg - 1) // agg > 256:
        agg *= 2
    return agg


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


def input_features(X, row_ptr, near_ptr, cmask, agg):
    """The 16 input features per point (X, the mask and the rows in internal order)."""
    n = X.shape[0]
    lo, hi = X.min(0), X.max(0)
    half = np.maximum(0.5 * (hi - lo), 1e-300)
    Xh = (X - 0.5 * (hi + lo)) / half
    deg = np.diff(near_ptr).astype(np.float64)
    cnt = np.diff(row_ptr).astype(np.float64)
    rel = np.zeros_like(X)
    for I in range((n + agg - 1) // agg):
        sl = slice(I * agg, min(n, (I + 1) * agg))
        cen = X[sl].mean(0)
        h = max(np.abs(X[sl] - cen).max(), 1e-300)
        rel[sl] = (X[sl] - cen) / h
    Fi = np.zeros((n, F_IN))
    Fi[:, 0:3] = Xh
    Fi[:, 3:6] = rel
    Fi[:, 6:9] = cmask
    Fi[:, 9] = deg / max(deg.max(), 1.0)
    Fi[:, 10] = cnt / max(cnt.max(), 1.0)
    Fi[:, 11:14] = rel ** 2
    Fi[:, 14] = 1.0
    return Fi


def apply_dinv(dinv, y):
    """Block Jacobi: q_i = R_i y_i, R_i stored column-major 3x3 (9 per point)."""
    R = dinv.reshape(-1, 3, 3)  # R[i, col, row]
    return np.einsum("icr,ic->ir", R, y.reshape(-1, 3))


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


def basis(O, cmask, use_fields=True):
    """B per point: (n, 3 dofs, 11 columns): translations, then the 8 fields, zero on
    constrained dofs."""
    n = O.shape[0]
    B = np.zeros((n, 3, 3 + K_F))
    for d in range(3):
        B[:, d, d] = 1.0
    if use_fields:
        B[:, :, 3:] = O.reshape(n, K_F, 3).transpose(0, 2, 1)
    return np.where(cmask[:, :, None], 0.0, B)


def lsq_loss(O, E, cmask, agg, use_fields=True):
    """The loss of the samples E (S, n, 3) for the fields O (n, 24) and its gradient dO."""
    n = O.shape[0]
    S = E.shape[0]
    free = ~cmask
    Bp = basis(O, cmask, use_fields)
    dB = np.zeros_like(Bp)
    w = np.array([1.0 / max(np.sum(np.where(free, e, 0.0) ** 2), 1e-300) for e in E]) / S
    total = 0.0
    nb = 3 + K_F
    for I in range((n + agg - 1) // agg):
        sl = slice(I * agg, min(n, (I + 1) * agg))
        B = Bp[sl].reshape(-1, nb)
        G = B.T @ B
        lam = RIDGE * np.trace(G) / nb + 1e-30
        A = G + lam * np.eye(nb)
        for s in range(S):
            e = np.where(cmask[sl], 0.0, E[s, sl]).reshape(-1)
            c = np.linalg.solve(A, B.T @ e)
            r = e - B @ c
            total += w[s] * (r @ r + lam * (c @ c))
            g = -2.0 * w[s] * np.outer(r, c) + w[s] * (2.0 * RIDGE / nb * (c @ c)) * B
            dB[sl] += g.reshape(-1, 3, nb)
    dB = np.where(cmask[:, :, None], 0.0, dB)
    dO = dB[:, :, 3:].transpose(0, 2, 1).reshape(n, N_OUT)
    return total, (dO if use_fields else np.zeros_like(dO))


def loss_and_grad(theta, E, graph, bf16_mode=False, want_grad=True):
    """E: error samples (S, n, 3); graph: dict with near_ptr, near, cmask (n, 3 bool), feat
    (n, 16), agg.  Returns (loss, grad)."""
    O, cache = forward(theta, graph["feat"], graph["near_ptr"], graph["near"], bf16_mode,
                       keep=True)
    loss, dO = lsq_loss(O, E, graph["cmask"], graph["agg"])
    if not want_grad:
        return loss, None
    return loss, backward(theta, cache, dO, graph["near_ptr"], graph["near"], bf16_mode)


def fields(theta, graph, bf16_mode=False):
    """The 8 fields per point, (n, 8, 3)."""
    O = forward(theta, graph["feat"], graph["near_ptr"], graph["near"], bf16_mode)
    return O.reshape(-1, K_F, 3)
