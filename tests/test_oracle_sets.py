"""The oracle's several-parameter-set path (a body whose points carry different material
parameters: material!(body, set_name; ...) per point set, ParameterHandler), pinned on
closed forms of two-point bodies, on the reduction to the single-set path when all sets are
equal, and on momentum conservation.

What a plausible mistake does here:
- using bc_i (or G_i, K_i) of the owner instead of the two ends' mean fails the two-point
  closed forms, whose sets differ by 50 %;
- using V_i for the term a neighbour scatters fails the volume-[1, 2] OSB closed form;
- failing a bond against the neighbour's eps_c fails the asymmetric-failure case;
- rho or bc of one global set in the time step or the relaxation density fails their
  per-point closed forms."""
import math

import numpy as np
import pytest

from conftest import as_oracle_params

DELTA = 1.5


def two_point_params(kind, E, epsc=0.0, rho=7850.0):
    import peridynamics as pd
    mat = {"BB": pd.BBMaterial, "OSB": pd.OSBMaterial, "CPD": pd.CMaterial}[kind]()
    kw = dict(horizon=DELTA, rho=rho, E=E)
    if kind != "BB":
        kw["nu"] = 0.25
    p = pd.materials.get_point_params(mat, kw)
    d = as_oracle_params(p, **(dict(model="SVK", Cs=0.5, maxdmg=0.85) if kind == "CPD" else {}))
    d["epsc"] = epsc
    return d


def two_points(oracle, kind, Ea, Eb, vol=(1.0, 1.0), epsc=(0.0, 0.0), rho=(7850.0, 7850.0)):
    X = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    sets = [two_point_params(kind, Ea, epsc[0], rho[0]), two_point_params(kind, Eb, epsc[1], rho[1])]
    fp = np.ones(2, np.uint8) if any(epsc) else np.zeros(2, np.uint8)
    return oracle.OracleSim(X, np.array(vol), kind, sets, fail_permit=fp, point_set=[0, 1]), sets


def test_bb_interface_closed_form(oracle):
    orc, (pa, pb) = two_points(oracle, "BB", 105e9, 70e9, vol=(1.0, 2.0))
    e = 0.0015
    x = orc.X.copy()
    x[1, 0] += e
    e = x[1, 0] - 1.0  # the stretch as rounded in the position (exact difference)
    b = orc.force(x)
    bc = (pa["bc"] + pb["bc"]) / 2
    # b_1 = bc_mean eps V_2 (along +x), b_2 = -bc_mean eps V_1, eps = e / L with L = 1
    np.testing.assert_allclose(b[0], [bc * e * 2.0, 0, 0], rtol=1e-14, atol=0)
    np.testing.assert_allclose(b[1], [-bc * e * 1.0, 0, 0], rtol=1e-14, atol=0)
    # K = E / (3 (1 - 2 nu)) with nu = 1/4, bc = 18 K / (pi delta^4)
    assert bc == pytest.approx(18 * ((105e9 + 70e9) / 2 / 1.5) / (math.pi * DELTA ** 4), rel=1e-14)


def test_osb_interface_closed_form(oracle):
    orc, (pa, pb) = two_points(oracle, "OSB", 105e9, 70e9, vol=(1.0, 2.0))
    e = 0.0015
    x = orc.X.copy()
    x[1, 0] += e
    e = x[1, 0] - 1.0  # the stretch as rounded in the position (exact difference)
    b = orc.force(x)
    V1, V2, w = 1.0, 2.0, DELTA / 1.0  # linear kernel delta / L, L = 1
    Gs, Ks = pa["G"] + pb["G"], pa["K"] + pb["K"]

    def bond_force(wvol, dil):
        c1 = 15.0 * Gs / (2 * wvol)
        c2 = dil * (3.0 * Ks / (2 * wvol) - c1 / 3.0)
        return w * (c2 * 1.0 + c1 * e)  # times dx / l = 1 along x

    # point 1 sees V_2, point 2 sees V_1: weighted volume w L^2 V_j, dilatation 3 / wvol w L e V_j
    p12 = bond_force(w * V2, 3.0 / (w * V2) * w * e * V2)
    p21 = bond_force(w * V1, 3.0 / (w * V1) * w * e * V1)
    b1 = p12 * V2 + p21 * V2   # own term V_j, and -(-p21) V_2 scattered by point 2
    b2 = -p12 * V1 - p21 * V1
    np.testing.assert_allclose(b[0], [b1, 0, 0], rtol=1e-13, atol=0)
    np.testing.assert_allclose(b[1], [b2, 0, 0], rtol=1e-13, atol=0)
    assert abs(b[0, 0] * V1 + b[1, 0] * V2) <= 1e-15 * abs(b[0, 0])


def test_osb_interface_fixture(oracle, golden):
    """The reference's known answer across an OSB material interface
    (test_ordinary_state_based.jl:31-61): E_b = 2 E_a, volumes [1, 2]."""
    import peridynamics as pd
    g = golden["osb_interface"]
    X = np.array(g["position"]).T
    sets = []
    for E in (g["E_a"], g["E_b"]):
        p = pd.materials.get_point_params(pd.OSBMaterial(), dict(
            horizon=g["horizon"], rho=g["rho"], E=E, nu=g["nu"], Gc=g["Gc"]))
        sets.append(as_oracle_params(p))
    orc = oracle.OracleSim(X, np.array(g["volume"]), "OSB", sets, fail_permit=np.zeros(2),
                           point_set=[0, 1])
    x = orc.X.copy()
    x[1, 0] = g["moved_x_of_point_1"]
    b = orc.force(x)
    np.testing.assert_allclose(b[0], [g["b1"], 0, 0], rtol=1e-13, atol=0)
    np.testing.assert_allclose(b[1], [-g["b2"], 0, 0], rtol=1e-13, atol=0)


def test_osb_asymmetric_failure(oracle):
    """eps_c of point 1 below the stretch, of point 2 above: the bond 1 -> 2 breaks and 2 -> 1
    stays.  Point 1 then has a zero weighted volume and contributes nothing; point 2's term
    still acts on both points."""
    orc, (pa, pb) = two_points(oracle, "OSB", 105e9, 70e9, vol=(1.0, 2.0), epsc=(1e-3, 1e-2))
    e = 0.0015
    x = orc.X.copy()
    x[1, 0] += e
    e = x[1, 0] - 1.0  # the stretch as rounded in the position (exact difference)
    b = orc.force(x)
    assert list(orc.active) == [0, 1]
    V1, V2, w = 1.0, 2.0, DELTA
    Gs, Ks = pa["G"] + pb["G"], pa["K"] + pb["K"]
    wvol = w * V1
    dil = 3.0 / wvol * w * e * V1
    c1 = 15.0 * Gs / (2 * wvol)
    c2 = dil * (3.0 * Ks / (2 * wvol) - c1 / 3.0)
    p21 = w * (c2 + c1 * e)
    np.testing.assert_allclose(b[0], [p21 * V2, 0, 0], rtol=1e-13, atol=0)
    np.testing.assert_allclose(b[1], [-p21 * V1, 0, 0], rtol=1e-13, atol=0)
    assert list(orc.damage) == [1.0, 0.0]


def test_bb_failure_uses_owner_eps_c(oracle):
    orc, _ = two_points(oracle, "BB", 105e9, 70e9, epsc=(1e-3, 1e-2))
    x = orc.X.copy()
    x[1, 0] += 0.0015
    b = orc.force(x)
    assert list(orc.active) == [0, 1]
    assert b[0, 0] == 0.0 and b[1, 0] < 0.0


def test_timestep_and_relaxation_density_per_point(oracle):
    orc, (pa, pb) = two_points(oracle, "BB", 105e9, 70e9, vol=(1.0, 2.0), rho=(7850.0, 2700.0))
    # dt_i = sqrt(2 rho_i / (V_j bc_i / L)), L = 1
    dt1 = math.sqrt(2 * 7850.0 / (2.0 * pa["bc"]))
    dt2 = math.sqrt(2 * 2700.0 / (1.0 * pb["bc"]))
    assert orc.stable_dt(1.0) == pytest.approx(min(dt1, dt2), rel=1e-15)
    lam, dt = 1.0, 1.0
    dens = orc.dr_density(dt, lam)
    k = 5 * math.pi * DELTA ** 2
    np.testing.assert_allclose(dens, [lam / 4 * dt * dt * 1 * k * pa["bc"],
                                      lam / 4 * dt * dt * 1 * k * pb["bc"]], rtol=1e-15)


@pytest.mark.parametrize("kind", ["BB", "OSB", "CPD"])
def test_equal_sets_reduce_to_single_set(oracle, kind):
    import peridynamics as pd
    pos, vol = pd.lattice_block(9, 8, 5, 1e-3, jitter_amplitude=0.05, seed=5)
    pos = np.ascontiguousarray(pos.T)  # (3, n) -> (n, 3)
    p = two_point_params(kind, 210e9)
    p["delta"] = 3.015e-3
    p["epsc"] = 2e-4
    rng = np.random.default_rng(0)
    ps = rng.integers(0, 3, len(vol))
    one = oracle.OracleSim(pos, vol, kind, p)
    many = oracle.OracleSim(pos, vol, kind, [p, dict(p), dict(p)], point_set=ps)
    x = pos + 4e-7 * rng.standard_normal(pos.shape)
    for _ in range(2):
        b1 = one.force(x).copy()
        b2 = many.force(x).copy()
        assert np.array_equal(b1, b2)
        assert np.array_equal(one.active, many.active)
        x = x + 3e-7 * rng.standard_normal(pos.shape)


@pytest.mark.parametrize("kind", ["BB", "OSB", "CPD"])
def test_momentum_conservation_sets(oracle, kind):
    """sum_i V_i b_i = 0 for any deformation.  OSB and CPD scatter each bond's term to both
    ends, so this holds with failures against differing eps_c too; BB gathers each point's
    own bonds, so a bond broken at one end only leaves an unbalanced force there (as in the
    reference), and its sets share eps_c here."""
    import peridynamics as pd
    pos, vol = pd.lattice_block(10, 9, 6, 1e-3, jitter_amplitude=0.05, seed=7)
    pos = np.ascontiguousarray(pos.T)  # (3, n) -> (n, 3)
    vol = vol * np.random.default_rng(1).uniform(0.7, 1.3, len(vol))
    sets = []
    for E, epsc in ((210e9, 3e-4), (70e9, 1e-3), (140e9, 6e-4)):
        p = two_point_params(kind, E)
        p["delta"] = 3.015e-3
        p["epsc"] = epsc if kind != "BB" else 6e-4
        sets.append(p)
    ps = (pos[:, 0] > np.median(pos[:, 0])).astype(int) + (pos[:, 1] > np.median(pos[:, 1]))
    orc = oracle.OracleSim(pos, vol, kind, sets, point_set=ps)
    x = pos + 6e-7 * np.random.default_rng(2).standard_normal(pos.shape)
    b = orc.force(x)
    assert 0 < orc.active.sum() < len(orc.active)  # some bonds broke
    total = (b * orc.vol[:, None]).sum(0)
    scale = (np.abs(b) * orc.vol[:, None]).sum(0)
    assert np.all(np.abs(total) <= 1e-12 * scale)


def test_cpd_sets_use_the_params_of_the_point(oracle):
    """CMaterial with several sets (force_density_point! receives the params of the point
    whose loop runs, correspondence.jl:183-191): a point whose neighbours all share its set
    feels exactly what a body of that one set gives it, on each side of the interface.  A
    point-independent choice (the first set, or bond-averaged constants as BB and OSB use)
    fails on one side."""
    import peridynamics as pd
    pos, vol = pd.lattice_block(14, 8, 6, 1e-3, jitter_amplitude=0.05, seed=3)
    pos = np.ascontiguousarray(pos.T)
    sets = []
    for E in (210e9, 70e9):
        p = two_point_params("CPD", E)
        p["delta"] = 3.015e-3
        sets.append(p)
    ps = (pos[:, 0] > np.median(pos[:, 0])).astype(int)
    many = oracle.OracleSim(pos, vol, "CPD", sets, fail_permit=np.zeros(len(vol)), point_set=ps)
    x = pos + 5e-7 * np.random.default_rng(4).standard_normal(pos.shape)
    b = many.force(x).copy()
    owner = np.repeat(np.arange(len(vol)), np.diff(many.row_ptr))
    mixed_family = np.zeros(len(vol), bool)
    np.logical_or.at(mixed_family, owner, ps[owner] != ps[many.nbr])
    checked = 0
    for k in (0, 1):
        one = oracle.OracleSim(pos, vol, "CPD", sets[k], fail_permit=np.zeros(len(vol)))
        bk = one.force(x)
        # the neighbours' neighbours matter too (their Q enters the reverse terms)
        far = ~mixed_family & (ps == k)
        far &= ~np.isin(np.arange(len(vol)), many.nbr[mixed_family[owner]])
        assert far.any()
        scale = np.max(one.b_abs, axis=1)
        assert np.all(np.abs(b[far] - bk[far]).max(1) <= 1e-12 * scale[far])
        checked += int(far.sum())
    assert checked > 50
    # and the interface points differ from both single-set bodies
    assert not np.allclose(b[mixed_family], bk[mixed_family])
