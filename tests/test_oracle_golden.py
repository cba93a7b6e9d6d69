"""Pins the CPU oracle against the reference's own known answers, invariants and
closed forms (no GPU).  Each test cites the reference test it restates."""
import math

import numpy as np
import pytest

from conftest import as_oracle_params


def _params(mat, **kw):
    import peridynamics as pd
    return pd.materials.get_point_params(mat, kw)


def _bb(**kw):
    import peridynamics as pd
    return _params(pd.BBMaterial(), **kw)


# --------------------------------------------------------------------- bond search
def test_tetra_bond_list_exact(oracle, golden):
    g = golden["tetra_bonds"]
    X = np.array(g["position"]).T
    for brute in (False, True):
        row_ptr, nbr, L, bfp = oracle.find_bonds(X, g["horizon"], brute=brute)
        assert list(np.diff(row_ptr)) == g["n_neighbors"]
        got = [[int(j), float(l)] for j, l in zip(nbr, L)]
        assert got == g["bonds"]  # exact: neighbour order and bit-exact lengths


def test_two_point_cases(oracle, golden):
    g = golden["find_bonds_two_points"]
    X = np.array(g["position"]).T
    for case in g["cases"]:
        row_ptr, nbr, L, bfp = oracle.find_bonds(X, case["delta"], np.array(case["fail_permit"]))
        got = [[int(nbr[b]), float(L[b]), bool(bfp[b])] for b in range(row_ptr[0], row_ptr[1])]
        assert got == case["bonds_of_0"]


def test_box_4x4x4_bond_count(oracle, golden):
    import peridynamics as pd
    g = golden["box_4x4x4_bonds"]
    pos, vol = pd.uniform_box(*g["uniform_box"])
    row_ptr, nbr, L, bfp = oracle.find_bonds(pos.T, g["horizon"])
    assert pos.shape[1] == g["n_points"]
    assert row_ptr[-1] == g["n_bonds"]
    assert np.all(np.diff(row_ptr) == g["n_active_per_point"])


def test_cell_list_equals_brute_force(oracle):
    import peridynamics as pd
    pos, _ = pd.lattice_block(9, 8, 7, 1e-3, jitter_amplitude=0.05, seed=7)
    a = oracle.find_bonds(pos.T, 3.015e-3)
    b = oracle.find_bonds(pos.T, 3.015e-3, brute=True)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    # symmetry: j in N(i) <=> i in N(j), with the same length bit for bit
    row_ptr, nbr, L, _ = a
    owner = np.repeat(np.arange(len(row_ptr) - 1), np.diff(row_ptr))
    fwd = {(int(i), int(j)): l for i, j, l in zip(owner, nbr, L)}
    assert all(fwd[(j, i)] == l for (i, j), l in fwd.items())


def test_lattice_neighbour_count(oracle):
    # an interior point of a cubic lattice has 122 neighbours within 3.015 dx
    import peridynamics as pd
    pos, _ = pd.lattice_block(9, 9, 9, 1.0)
    row_ptr, *_ = oracle.find_bonds(pos.T, 3.015)
    centre = 4 + 9 * (4 + 9 * 4)
    assert row_ptr[centre + 1] - row_ptr[centre] == 122


def test_duplicate_point_raises(oracle):
    X = np.array([[0.0, 0, 0], [1.0, 0, 0], [1.0, 0, 0]])
    with pytest.raises(oracle.DuplicatePointError):
        oracle.find_bonds(X, 1.5)


def test_precrack_hits_only_crossing_bonds(oracle):
    import peridynamics as pd
    pos, _ = pd.lattice_block(6, 6, 2, 1.0)
    X = pos.T
    row_ptr, nbr, L, _ = oracle.find_bonds(X, 3.015)
    a = np.where(X[:, 1] > 0)[0]
    b = np.where(X[:, 1] < 0)[0]
    ina = np.zeros(len(X), np.uint8); ina[a] = 1
    inb = np.zeros(len(X), np.uint8); inb[b] = 1
    hit = oracle.precrack_hits(row_ptr, nbr, ina, inb)
    owner = np.repeat(np.arange(len(X)), np.diff(row_ptr))
    crossing = (X[owner, 1] > 0) != (X[nbr, 1] > 0)
    assert np.array_equal(hit, crossing)


# --------------------------------------------------------------------- force densities
def test_bb_stretched_bond_closed_form(oracle, golden):
    g = golden["bb_stretched_bond"]
    X = np.array(g["position"]).T
    p = _bb(horizon=g["horizon"], rho=g["rho"], E=g["E"])
    sim = oracle.OracleSim(X, g["volume"], "BB", as_oracle_params(p), fail_permit=[0, 0])
    x = X.copy()
    x[1, 0] = g["moved_x_of_point_1"]
    b = sim.force(x)
    E, d = g["E"], g["horizon"]
    b12 = 18 * E / (3 * (1 - 2 * 0.25)) / (math.pi * d ** 4) * 1.0015 * 0.0015 / 1.0015 * 1.0
    np.testing.assert_allclose(b[0], [b12, 0, 0], rtol=1e-13, atol=0)
    np.testing.assert_allclose(b[1], [-b12, 0, 0], rtol=1e-13, atol=0)


def test_osb_stretched_bond_known_answer(oracle, golden):
    import peridynamics as pd
    g = golden["osb_stretched_bond"]
    X = np.array(g["position"]).T
    p = _params(pd.OSBMaterial(), horizon=g["horizon"], rho=g["rho"], E=g["E"], nu=g["nu"],
                Gc=g["Gc"])
    sim = oracle.OracleSim(X, g["volume"], "OSB", as_oracle_params(p), fail_permit=[0, 0])
    x = X.copy()
    x[1, 0] = g["moved_x_of_point_1"]
    b = sim.force(x)
    np.testing.assert_allclose(b[0], [g["b12"], 0, 0], rtol=1e-13)
    np.testing.assert_allclose(b[1], [-g["b12"], 0, 0], rtol=1e-13)


def test_cpd_nhp_regression_vector(oracle, golden):
    import peridynamics as pd
    g = golden["cpd_nhp_displaced_point"]
    X = np.array(g["position"]).T
    p = _params(pd.CMaterial(model=pd.NeoHookePenalty()), horizon=g["horizon"], rho=g["rho"],
                E=g["E"], nu=g["nu"], Gc=g["Gc"])
    sim = oracle.OracleSim(X, g["volume"], "CPD", as_oracle_params(p, model="NHP"),
                           fail_permit=np.zeros(5))
    x = X.copy()
    x[1, 0] += g["shift_x_of_point_1"]
    b = sim.force(x)
    assert np.all(b[4] == 0)  # the isolated point feels nothing
    # The reference asserts these at rtol = sqrt(eps).  The restatement agrees to 1e-11 of the
    # force scale; the remaining difference is the op order of StaticArrays' inv/det (the
    # y component of point 2 is a difference of nearly equal terms).
    tol = dict(rtol=0, atol=1e-11 * np.abs(b).max())
    np.testing.assert_allclose(b[0], g["b_int_point_0"], **tol)
    np.testing.assert_allclose(b[1, 0], g["b_int_x_point_1"], **tol)
    np.testing.assert_allclose(b[2, 1], g["b_int_y_point_2"], **tol)
    assert np.all(np.abs(b.sum(0)) < 1e-12 * np.abs(b).max())


def _cube(kind, n=6, m=3.015, model="SVK", Gc=2.7):
    """Fixtures.cube: unit cube of n^3 points, steel-like, no_failure! (fixtures.jl:50-58)."""
    import peridynamics as pd
    dx = 1.0 / n
    pos, vol = pd.uniform_box(1.0, 1.0, 1.0, dx)
    mat = {"BB": pd.BBMaterial(), "OSB": pd.OSBMaterial(), "CPD": pd.CMaterial()}[kind]
    p = _params(mat, horizon=m * dx, rho=7850.0, E=210e9, nu=0.25, Gc=Gc)
    return pos.T.copy(), vol, as_oracle_params(p, model=model)


def _deform(X, dt, F=((1.01, 0.003, 0.0), (0.0, 0.995, 0.0), (0.0, 0.0, 1.0)), amp=1e-4):
    """Fixtures.deform!: F X plus a sinpi perturbation (fixtures.jl:123-141)."""
    F = np.array(F)
    x = X @ F.T
    x[:, 0] += amp * np.sin(np.pi * 3 * X[:, 1])
    x[:, 1] += amp * np.sin(np.pi * 3 * X[:, 2])
    x[:, 2] += amp * np.sin(np.pi * 3 * X[:, 0])
    return x


KINDS = ["BB", "OSB", "CPD"]


F_AFFINE = np.array([[1.1, 0.01, 0.0], [0.0, 1.2, 0.02], [0.03, 0.0, 1.3]])


@pytest.mark.parametrize("model", ["SVK", "NHP"])
def test_cpd_defgrad_exact_for_affine_deformation(oracle, model):
    """test/materials/test_uniform_deformation.jl:5-41: under x = F X the deformation
    gradient of every point, surface points with truncated families included, is F (the
    nonlocal average of an affine map is the map).  A dropped bond term, a transposed F or
    a wrong K^-1 product breaks it at the surface."""
    X, vol, p = _cube("CPD", n=5, m=3.1, model=model)
    sim = oracle.OracleSim(X, vol, "CPD", p, fail_permit=np.zeros(len(X)))
    sim.force(X @ F_AFFINE.T)
    F_cm = F_AFFINE.T.reshape(-1)  # column-major, F[r + 3c]
    assert np.max(np.abs(sim.defgrad - F_cm)) < 1e-12


@pytest.mark.parametrize("kind", KINDS)
def test_zero_force_at_reference(oracle, kind):
    # test/materials/test_force_density.jl:5-21
    X, vol, p = _cube(kind)
    sim = oracle.OracleSim(X, vol, kind, p, fail_permit=np.zeros(len(X)))
    b = sim.force()
    assert np.max(np.abs(b)) < 1e-12 * p["E"] * 6


@pytest.mark.parametrize("kind", KINDS)
def test_momentum_conservation(oracle, kind):
    # test/materials/test_force_density.jl:23-40
    X, vol, p = _cube(kind)
    sim = oracle.OracleSim(X, vol, kind, p, fail_permit=np.zeros(len(X)))
    b = sim.force(_deform(X, 1.0))
    total = (b * vol[:, None]).sum(0)
    scale = (np.abs(b) * vol[:, None]).sum(0)
    assert np.all(scale > 0)
    assert np.all(np.abs(total) < 1e-10 * scale)


@pytest.mark.parametrize("kind", KINDS)
def test_momentum_conservation_nonuniform_volume(oracle, kind):
    # the same balance with point volumes varying by +-30 %: it holds only if the term a
    # point scatters into its neighbour carries the volume of the emitting point
    # (ordinary_state_based.jl:145-146, correspondence.jl:248-249)
    X, vol, p = _cube(kind)
    vol = vol * np.random.default_rng(5).uniform(0.7, 1.3, len(vol))
    sim = oracle.OracleSim(X, vol, kind, p, fail_permit=np.zeros(len(X)))
    b = sim.force(_deform(X, 1.0))
    total = (b * vol[:, None]).sum(0)
    scale = (np.abs(b) * vol[:, None]).sum(0)
    assert np.all(np.abs(total) < 1e-10 * scale)


@pytest.mark.parametrize("kind", KINDS)
def test_mirror_symmetry(oracle, kind):
    # test/materials/test_force_density.jl:42-84
    X, vol, p = _cube(kind)
    sim = oracle.OracleSim(X, vol, kind, p, fail_permit=np.zeros(len(X)))
    F = np.diag([1.01, 0.995, 1.02])
    b = sim.force(X @ F.T)
    assert np.any(b != 0)
    for d in range(3):
        sgn = np.ones(3); sgn[d] = -1
        Xm = X * sgn
        mirror = np.array([np.argmin(np.abs(X - xm).sum(1)) for xm in Xm])
        assert np.allclose(X[mirror], Xm, atol=1e-12)
        for e in range(3):
            s = -1 if d == e else 1
            np.testing.assert_allclose(b[:, e], s * b[mirror, e], atol=1e-10 * np.abs(b).max(),
                                       rtol=0)


@pytest.mark.parametrize("kind", KINDS)
def test_pulled_corner_breaks_all_its_bonds(oracle, kind):
    # test/materials/test_force_density.jl:86-124
    import peridynamics as pd
    pos, vol = pd.uniform_box(1, 1, 1, 1 / 3)
    mat = {"BB": pd.BBMaterial(), "OSB": pd.OSBMaterial(), "CPD": pd.CMaterial()}[kind]
    p = _params(mat, horizon=3.015 / 3, rho=7850.0, E=210e9, nu=0.25, Gc=1e-6)
    X = pos.T.copy()
    sim = oracle.OracleSim(X, vol, kind, as_oracle_params(p), fail_permit=np.ones(len(X)))
    x = X.copy()
    x[0] += 2.0
    sim.force(x)
    assert sim.damage[0] == pytest.approx(1.0)
    assert sim.n_active[0] == 0
    for i in range(1, len(X)):
        nb = sim.nbr[sim.row_ptr[i]:sim.row_ptr[i + 1]]
        lost = int(np.sum(nb == 0))
        if kind == "CPD" and sim.damage[i] > 0.85:
            continue
        assert sim.n_active[i] == len(nb) - lost
        assert sim.damage[i] == pytest.approx(lost / len(nb))


def test_stable_timestep_closed_form(oracle, golden):
    g = golden["stable_timestep_two_points"]
    X = np.array(g["position"]).T
    p = _bb(horizon=g["horizon"], rho=g["rho"], E=g["E"], Gc=g["Gc"])
    sim = oracle.OracleSim(X, g["volume"], "BB", as_oracle_params(p))
    E, nu, d, rho = g["E"], g["nu"], g["horizon"], g["rho"]
    bc = 18 * E / (3 * (1 - 2 * nu)) / (math.pi * d ** 4)
    assert abs(sim.stable_dt(0.7) - 0.7 * math.sqrt(2 * rho / bc)) <= 2.3e-16


# --------------------------------------------------------------------- time integration
def test_velocity_verlet_update_algebra(oracle):
    # test/unit/time_solvers/test_velocity_verlet.jl:112-250: v_half = v + a dt/2,
    # u += v_half dt, x += v_half dt, a = (b_int + b_ext)/rho, v = v_half + a dt/2
    lib = oracle.lib()
    v = np.array([[1.0, 3.0, 5.0], [2.0, 4.0, 6.0]])
    a = np.array([[0.1, 0.3, 0.5], [0.2, 0.4, 0.6]])
    vh = np.zeros((2, 3))
    lib.orc_update_vel_half(6, oracle._p(v), oracle._p(a), 0.5, oracle._p(vh))
    assert np.array_equal(vh, v + a * 0.5)
    u = np.ones((2, 3))
    x = np.full((2, 3), 2.0)
    lib.orc_update_disp_and_pos(6, oracle._p(vh), 0.25, oracle._p(u), oracle._p(x))
    assert np.array_equal(u, 1.0 + vh * 0.25)
    assert np.array_equal(x, 2.0 + vh * 0.25)
    bi = np.array([[10.0, 20, 30], [40, 50, 60]])
    be = np.array([[1.0, 2, 3], [4, 5, 6]])
    acc = np.zeros((2, 3)); vel = np.zeros((2, 3))
    lib.orc_update_acc_and_vel(6, oracle._p(bi), oracle._p(be), 2.0, oracle._p(vh), 0.5,
                               oracle._p(acc), oracle._p(vel))
    assert np.array_equal(acc, (bi + be) / 2.0)
    assert np.array_equal(vel, vh + acc * 0.5)


def test_vv_wave_speed_sign(oracle):
    # a body pulled apart in y by the configuration-1 velocity BCs develops tension: the
    # top row moves up, the bottom row down, and the momentum of the free points is balanced
    import peridynamics as pd
    pos, vol = pd.lattice_block(12, 12, 4, 1e-3)
    p = _bb(horizon=3.015e-3, rho=2700.0, E=70e9, Gc=100.0)
    X = pos.T.copy()
    sim = oracle.OracleSim(X, vol, "BB", as_oracle_params(p))
    dt = sim.stable_dt(0.7)
    y = X[:, 1]
    top = np.where(y > y.max() - 3e-3 + 1e-9)[0]
    bot = np.where(y < y.min() + 3e-3 - 1e-9)[0]
    bcs = [("sd", top, 1, lambda t: 0.6), ("sd", bot, 1, lambda t: -0.6)]
    sim.vv_run(20, dt, bcs)
    assert np.all(sim.u[top, 1] > 0) and np.all(sim.u[bot, 1] < 0)
    np.testing.assert_allclose(sim.u[top, 1], 20 * 0.6 * dt, rtol=1e-12)


# --------------------------------------------------------------------- Newton-Krylov
def test_nk_residual(oracle, golden):
    # test/unit/time_solvers/test_newton_krylov.jl:330-362
    g = golden["nk_residual"]
    X = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    p = _bb(horizon=2, rho=1, E=1, Gc=None) if False else _bb(horizon=2, rho=1, E=1)
    sim = oracle.OracleSim(X, g["volume"], "BB", as_oracle_params(p), fail_permit=np.zeros(4))
    sim.b_int[0, 0] = g["b_int_x_point_0"]
    sim.b_ext[0, 0] = g["b_ext_x_point_0"]
    r = sim.residual(constrained=np.array([3 * 1 + 1]))
    assert r[0] == pytest.approx(g["residual_x_point_0"])
    assert r[4] == 0.0


@pytest.mark.slow
def test_nk_uniform_tension_closed_form(oracle):
    # test/simulation/test_analytic_solutions.jl:7-94: a CMaterial bar of 297 points pulled by
    # F over its right face converges to the elongation F l / (E A) within 2 %.
    import peridynamics as pd
    l, w, h, Ftot, E = 1.0, 0.1, 0.1, 2e6, 200e9
    dx = 1 / 30
    pos, vol = pd.uniform_box(l + 3 * dx, w, h, dx, center=(-1.5 * dx, 0, 0))
    X = pos.T.copy()
    p = _params(pd.CMaterial(), horizon=3.015 * dx, rho=7850.0, E=E, nu=0.25)
    sim = oracle.OracleSim(X, vol, "CPD", as_oracle_params(p, model="SVK", Cs=0.5),
                           fail_permit=np.zeros(len(X)))
    left = np.where(X[:, 0] < -l / 2)[0]
    right = np.where(X[:, 0] > l / 2 - dx)[0]
    b_right = Ftot / vol[right].sum()
    sim.b_ext[right, 0] = b_right
    disp = [(left, d, lambda p: 0.0) for d in range(3)]
    for n in range(1, 6):
        # forcedensity_bc!(p -> b_right) scales with the load factor like the displacements
        sim.b_ext[right, 0] = n / 5 * b_right
        sim.nk_step(n, 5, disp, maxiter=50, tol=1e-3)
    elong = sim.u[right, 0].mean()
    exact = Ftot / (E * w * h) * l
    assert abs(elong - exact) / exact < 0.02
