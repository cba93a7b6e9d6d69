"""The oracle's constitutive models (first_piola_kirchhoff of constitutive_models.jl:33-255:
SaintVenantKirchhoff, LinearElastic, NeoHooke, NeoHookePenalty), pinned as the reference's
test/unit/physics/test_constitutive_models.jl pins them: zero stress at F = I, P = dPsi/dF
(here by central differences of the textbook strain energies, the reference uses
ForwardDiff), a symmetric Cauchy stress, positive stress under tension, LinearElastic equal
to Saint Venant-Kirchhoff, and the closed form of pure shear.

A transposed product (S F instead of F S), a dropped term of S or the wrong Voigt slot for a
shear component fails the energy-derivative check."""
import numpy as np
import pytest

E_MOD, NU = 210e9, 0.3
LAM = E_MOD * NU / ((1 + NU) * (1 - 2 * NU))
MU = E_MOD / (2 * (1 + NU))
K_MOD = E_MOD / (3 * (1 - 2 * NU))
F_TEST = np.array([[1.02, 0.013, -0.004], [0.006, 0.985, 0.011], [-0.009, 0.007, 1.017]])


def psi(model, F):
    """strain_energy_density of each model (constitutive_models.jl)."""
    C = F.T @ F
    J = np.linalg.det(F)
    if model in ("SVK", "LE"):
        E = 0.5 * (C - np.eye(3))
        return 0.5 * LAM * np.trace(E) ** 2 + MU * np.trace(E @ E)
    if model == "NH":
        return 0.5 * MU * (np.trace(C) - 3) - MU * np.log(J) + 0.5 * LAM * np.log(J) ** 2
    return 0.5 * MU * (np.trace(C) * J ** (-2 / 3) - 3) + K_MOD / 8 * (J ** 2 + J ** -2 - 2)


def P_of(oracle, model, F):
    return oracle.pk1(model, F, LAM, MU, MU, K_MOD, NU)


MODELS = ["SVK", "LE", "NH", "NHP"]


@pytest.mark.parametrize("model", MODELS)
def test_zero_stress_at_identity(oracle, model):
    assert np.max(np.abs(P_of(oracle, model, np.eye(3)))) <= 1e-6  # Pa, of ~1e11 moduli


@pytest.mark.parametrize("model", MODELS)
def test_pk1_is_the_energy_derivative(oracle, model):
    P = P_of(oracle, model, F_TEST)
    h = 1e-6
    dpsi = np.zeros((3, 3))
    for r in range(3):
        for c in range(3):
            Fp, Fm = F_TEST.copy(), F_TEST.copy()
            Fp[r, c] += h
            Fm[r, c] -= h
            dpsi[r, c] = (psi(model, Fp) - psi(model, Fm)) / (2 * h)
    np.testing.assert_allclose(P, dpsi, rtol=0, atol=1e-6 * np.abs(P).max())


@pytest.mark.parametrize("model", MODELS)
def test_cauchy_symmetric_and_tension_positive(oracle, model):
    P = P_of(oracle, model, F_TEST)
    sig = P @ F_TEST.T / np.linalg.det(F_TEST)
    np.testing.assert_allclose(sig, sig.T, rtol=0, atol=1e-12 * np.abs(sig).max())
    Pt = P_of(oracle, model, np.diag([1.001, 1.0, 1.0]))
    assert Pt[0, 0] > 0


def test_linear_elastic_equals_svk(oracle):
    np.testing.assert_allclose(P_of(oracle, "LE", F_TEST), P_of(oracle, "SVK", F_TEST),
                               rtol=1e-13, atol=0)


@pytest.mark.parametrize("model", ["SVK", "LE"])
def test_pure_shear_closed_form(oracle, model):
    """F = [1 g 0; 0 1 0; 0 0 1]: E = [0 g/2 0; g/2 g^2/2 0; 0 0 0], S = lam tr(E) I + 2 mu E.
    E22 = (1 + g^2 - 1) / 2 carries the rounding of 1 + g^2 (relative ~1e-10 here), so the
    comparison is the reference's isapprox (rtol sqrt(eps))."""
    g = 1e-3
    F = np.array([[1.0, g, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    E = np.array([[0.0, g / 2, 0.0], [g / 2, g * g / 2, 0.0], [0.0, 0.0, 0.0]])
    S = LAM * np.trace(E) * np.eye(3) + 2 * MU * E
    np.testing.assert_allclose(P_of(oracle, model, F), F @ S, rtol=1.5e-8,
                               atol=1.5e-8 * np.abs(F @ S).max())


# ZEMWan: calc_zem_stiffness_tensor (zem_stabilization.jl:32-42) and c_force_density!
# (correspondence.jl:253-279)

def _spd(seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((3, 3))
    return A @ A.T + 3 * np.eye(3)


def test_zem_c1_isotropic_closed_form(oracle):
    """For consistent isotropic constants a = (1 - nu) lambda / nu = lambda + 2 mu, and the
    contraction of the isotropic tensor with K^-1 is lambda tr(K^-1) I + 2 mu sym(K^-1)."""
    o = oracle
    for seed in range(4):
        Kinv = np.linalg.inv(_spd(seed))
        C1 = o.zem_c1(Kinv, LAM, MU, NU)
        want = LAM * np.trace(Kinv) * np.eye(3) + MU * (Kinv + Kinv.T)
        np.testing.assert_allclose(C1, want, rtol=0, atol=1e-13 * np.abs(want).max())


def test_zem_c1_brute_force_hooke_tensor(oracle):
    """C_1[i,j] = sum_kl C[i,j,k,l] Kinv[k,l] with the full 81-entry Hooke tensor built from
    its Voigt matrix, for constants where a differs from lambda + 2 mu (nu not the one that
    links lambda and mu) and a non-symmetric Kinv."""
    o = oracle
    lam, mu, nu = 1.3, 0.7, 0.21
    a = (1 - nu) * lam / nu
    Cv = np.zeros((6, 6))
    Cv[:3, :3] = lam
    Cv[[0, 1, 2], [0, 1, 2]] = a
    Cv[3, 3] = Cv[4, 4] = Cv[5, 5] = mu
    vm = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    C = np.zeros((3, 3, 3, 3))
    for I, (i1, i2) in enumerate(vm):
        for J, (j1, j2) in enumerate(vm):
            for (p, q) in ((i1, i2), (i2, i1)):
                for (r, s) in ((j1, j2), (j2, j1)):
                    C[p, q, r, s] = Cv[I, J]
    Kinv = np.random.default_rng(7).standard_normal((3, 3))
    want = np.einsum("ijkl,kl->ij", C, Kinv)
    np.testing.assert_allclose(o.zem_c1(Kinv, lam, mu, nu), want, rtol=0, atol=1e-14 * np.abs(want).max())


def _cparams(mat, horizon, **kw):
    import peridynamics as pd
    from conftest import as_oracle_params
    p = pd.materials.get_point_params(mat, dict(horizon=horizon, rho=1.0, E=1.0, nu=0.25, Gc=1.0))
    return as_oracle_params(p, **kw)


def test_zem_wan_equals_silling_on_the_tetrahedron(oracle, golden):
    """test/unit/physics/test_correspondence.jl:56-67: every family of the tetrahedron is
    complete, so z = dx - F dX vanishes on each bond and ZEMWan, ZEMSilling and no
    stabilisation give the same force density."""
    import peridynamics as pd
    g = golden["cpd_nhp_displaced_point"]
    X = np.array(g["position"]).T
    mat = pd.CMaterial(model=pd.NeoHookePenalty())
    out = {}
    for name, extra in (("silling", {}), ("wan", {"zem": "Wan"}), ("none", {"Cs": 0.0})):
        q = _cparams(mat, g["horizon"], model="NHP", **extra)
        sim = oracle.OracleSim(X, g["volume"], "CPD", q, fail_permit=np.zeros(5))
        x = X.copy()
        x[1, 0] += g["shift_x_of_point_1"]
        out[name] = sim.force(x).copy()
    scale = np.abs(out["silling"]).max()
    np.testing.assert_allclose(out["wan"], out["silling"], rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(out["none"], out["silling"], rtol=0, atol=1e-12 * scale)


def test_zem_wan_force_is_bare_force_plus_c1_z(oracle):
    """On a perturbed cube (z != 0) the ZEMWan force density minus the bare one (Cs = 0) is
    the sum over bonds of w C_1(i) z_ij V_j at i and its negative, times V_i, at j, with
    C_1 = lambda tr(K^-1) I + 2 mu sym(K^-1) from each point's shape tensor K = sum w V_j
    dX dX' and z_ij = dx - F_i dX.  Using C_1 of the neighbour, or dropping the reaction,
    fails it."""
    import peridynamics as pd
    n = 4
    pos, vol = pd.uniform_box(1.0, 1.0, 1.0, 1.0 / n)
    X = pos.T.copy()
    mat = pd.CMaterial()
    p = pd.materials.get_point_params(mat, dict(horizon=3.015 / n, rho=7850.0, E=210e9,
                                                nu=0.25, Gc=2.7))
    from conftest import as_oracle_params
    rng = np.random.default_rng(3)
    x = X @ np.array([[1.01, 0.003, 0.0], [0.0, 0.995, 0.0], [0.0, 0.0, 1.0]]).T
    x = x + 1e-4 * rng.standard_normal(x.shape) / n
    sims = {}
    for name, extra in (("wan", {"zem": "Wan"}), ("none", {"Cs": 0.0})):
        sim = oracle.OracleSim(X, vol, "CPD", as_oracle_params(p, model="SVK", **extra),
                               fail_permit=np.zeros(len(X)))
        sim.force(x)
        sims[name] = sim
    s = sims["wan"]
    want = np.zeros_like(X)
    for i in range(len(X)):
        js = s.nbr[s.row_ptr[i]:s.row_ptr[i + 1]]
        w = s.omega[s.row_ptr[i]:s.row_ptr[i + 1]] * s.active[s.row_ptr[i]:s.row_ptr[i + 1]]
        DX = X[js] - X[i]
        dx = x[js] - x[i]
        K = np.einsum("b,b,bi,bj->ij", w, vol[js], DX, DX)
        Kinv = np.linalg.inv(K)
        C1 = p.lam * np.trace(Kinv) * np.eye(3) + p.mu * (Kinv + Kinv.T)
        F = s.defgrad.reshape(-1)[9 * i:9 * i + 9].reshape(3, 3).T
        z = dx - DX @ F.T
        t = w[:, None] * (z @ C1.T)
        want[i] += (t * vol[js][:, None]).sum(0)
        np.add.at(want, js, -t * vol[i])
    got = sims["wan"].b_int - sims["none"].b_int
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-9 * np.abs(want).max())
    assert np.abs(want).max() > 1e-6 * np.abs(sims["none"].b_int).max()


@pytest.mark.parametrize("model", MODELS)
def test_stored_cauchy_stress_of_affine_deformation(oracle, model):
    """The cauchy_stress the oracle's CPD force stores per point (stress.jl:112-116), under
    the affine map x = F X: an interior point's F equals F_TEST to rounding, so its stored
    stress is P(F) F' / det F with P the energy derivative pinned above; the 9 entries of a
    point are a column-major 3x3 (as the reference's 9 x n field).  A point missing its outer
    bonds has the same F (a complete shape tensor is not needed for an affine map), so every
    point matches."""
    import peridynamics as pd
    pos, vol = pd.lattice_block(8, 8, 8, 1e-3)
    mat = pd.CMaterial(model={"SVK": pd.SaintVenantKirchhoff, "LE": pd.LinearElastic,
                              "NH": pd.NeoHooke, "NHP": pd.NeoHookePenalty}[model]())
    p = pd.materials.get_point_params(mat, dict(horizon=3.015e-3, rho=7850.0, E=E_MOD, nu=NU))
    from conftest import as_oracle_params
    X = pos.T.copy()
    sim = oracle.OracleSim(X, vol, "CPD", as_oracle_params(p, model=model),
                           fail_permit=np.zeros(len(X)))
    sim.force(X @ F_TEST.T)
    P = P_of(oracle, model, F_TEST)
    want = (P @ F_TEST.T / np.linalg.det(F_TEST)).reshape(-1, order="F")
    got = sim.cauchy
    assert got.shape == (len(X), 9)
    np.testing.assert_allclose(got, np.broadcast_to(want, got.shape), rtol=0,
                               atol=1e-9 * np.abs(want).max())
