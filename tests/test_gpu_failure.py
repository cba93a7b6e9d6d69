"""Bond failure, damage and CMaterial's too_much_damage! on the HIP path, against the oracle.

The pulled-corner case is the reference's test/materials/test_force_density.jl:86-124: a
corner point of a 3x3x3 cube is pulled far away with a tiny energy release rate, so all of
its bonds break in the first force evaluation; its damage is 1, and every former neighbour
lost exactly one bond.  For CMaterial the pulled point is also killed (damage > maxdmg,
correspondence.jl:288-298): n_active = 0 and every bond of the point inactive.

The kill paths of the CPD passes are exercised further by a chain of collinear points (a
singular shape tensor, so F holds NaN and every point is killed with damage 0) and by a
tearing bar with a low maxdmg, where points are killed during a Velocity Verlet run and
their neighbours must drop the killed points' terms.

Bars: bond states, n_active and damage bit-exact; FP64 force within 1e-12 of the oracle on
the per-point rounding scale (1e-5 in mixed precision); trajectories within 1e-9
(displacement) and 1e-8 (velocity) normwise; Cauchy stress within 1e-13 of the P-wave
modulus (an error in units of strain)."""
import sys

import numpy as np
import pytest

import peridynamics as pd
from conftest import as_oracle_params

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-5
STRESS_TOL = 1e-13
KINDS = ["BB", "OSB", "CPD"]
MATS = {"BB": pd.BBMaterial, "OSB": pd.OSBMaterial, "CPD": pd.CMaterial}
MODEL_NAMES = {pd.NeoHookePenalty: "NHP", pd.LinearElastic: "LE", pd.NeoHooke: "NH",
               pd.SaintVenantKirchhoff: "SVK"}


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def oracle_of(oracle, b, kind):
    """OracleSim of a body with one or several parameter sets."""
    extra = {}
    if kind == "CPD":
        extra = dict(model=MODEL_NAMES[type(b.mat.constitutive_model)], Cs=b.mat.zem.Cs,
                     maxdmg=b.mat.maxdmg)
    fp = b.fail_permit.astype(np.uint8)
    if len(b.point_params) > 1:
        sets = [as_oracle_params(p, **extra) for p in b.point_params]
        return oracle.OracleSim(b.xyz, b.volume, kind, sets, fail_permit=fp,
                                point_set=b.params_map - 1)
    return oracle.OracleSim(b.xyz, b.volume, kind, as_oracle_params(b.point_params[0], **extra),
                            fail_permit=fp)


def scaled_error(b_gpu, orc):
    scale = np.max(orc.b_abs, axis=1)
    diff = np.max(np.abs(b_gpu - orc.b_int), axis=1)
    m = scale > 0
    assert np.all(diff[~m] == 0)
    return float(np.max(diff[m] / scale[m])) if m.any() else 0.0


def normwise(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def states_equal(dh, orc):
    _, _, _, _, act = dh.bonds()
    assert np.array_equal(act, orc.active)
    np.testing.assert_array_equal(dh.get("n_active_bonds").astype(np.int64), orc.n_active)
    np.testing.assert_array_equal(dh.get("damage"), orc.damage)


# ------------------------------------------------------------------ pulled corner
def corner_body(kind, sets):
    pos, vol = pd.uniform_box(1, 1, 1, 1 / 3)
    b = pd.Body(MATS[kind](), pos, vol)
    kw = dict(horizon=3.015 / 3, rho=7850.0, E=210e9, nu=0.25, Gc=1e-6)
    if not sets:
        pd.material(b, **kw)
        return b
    # two sets split at z = 0, the second with a third of the stiffness
    pd.point_set(b, "low", pos[2] < 0)
    pd.point_set(b, "high", pos[2] >= 0)
    pd.material(b, "low", **kw)
    pd.material(b, "high", **dict(kw, E=70e9, rho=2700.0))
    return b


@pytest.mark.parametrize("sets", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kind", KINDS)
def test_pulled_corner(need_gpu, oracle, kind, precision, sets):
    say("pulled corner", kind, precision, "sets" if sets else "one set")
    b = corner_body(kind, sets)
    assert np.all(b.fail_permit)
    dh = pd.MI355XBodyDataHandler(b, precision=precision)
    orc = oracle_of(oracle, b, kind)
    pulled = 0
    x = b.xyz.copy()
    x[pulled] += 2.0  # far outside the cube
    dh.set("position", x)
    dh.calc_force_density(1e-7, 1e-7)
    orc.force(x)
    # the reference's own assertions, on the device's state
    dmg = dh.get("damage")
    nact = dh.get("n_active_bonds").astype(np.int64)
    row_ptr, nbr, _, _, act = dh.bonds()
    assert dmg[pulled] == pytest.approx(1.0)
    assert nact[pulled] == 0
    for i in range(1, b.n_points):
        nb = nbr[row_ptr[i]:row_ptr[i + 1]]
        lost = int(np.sum(nb == pulled))
        assert nact[i] == len(nb) - lost
        assert dmg[i] == pytest.approx(lost / len(nb))
    # and the oracle's state bit for bit
    states_equal(dh, orc)
    assert np.all(act[row_ptr[pulled]:row_ptr[pulled + 1]] == 0)
    err = scaled_error(dh.get("b_int"), orc)
    say(f"  scaled force error {err:.3e}")
    assert err <= (FP64_TOL if precision == "fp64" else MIXED_TOL)
    if kind == "CPD":
        # the pulled point is killed (too_much_damage!): it exerts and receives no force
        assert np.all(dh.get("b_int")[pulled] == 0)


# ------------------------------------------------------------------ CPD kill paths
def test_cpd_collinear_chain_is_killed(need_gpu, oracle):
    """Points on one line: the shape tensor K of every point is singular, F = Ft K^-1 holds
    NaN, and too_much_damage! kills every point although no bond is broken (damage 0,
    n_active 0); no force and no NaNError."""
    say("cpd collinear chain")
    n = 12
    pos = np.zeros((3, n))
    pos[0] = np.arange(n) * 1e-3
    b = pd.Body(pd.CMaterial(), pos, np.full(n, 1e-9))
    pd.material(b, horizon=3.015e-3, rho=7850.0, E=210e9, nu=0.25, Gc=100.0)
    dh = pd.MI355XBodyDataHandler(b)
    orc = oracle_of(oracle, b, "CPD")
    x = b.xyz.copy()
    x[:, 0] *= 1.0 + 1e-5
    dh.set("position", x)
    dh.calc_force_density(0.0, 1e-7)
    orc.force(x)
    assert np.all(orc.n_active == 0) and np.all(orc.damage == 0)
    states_equal(dh, orc)
    assert np.all(dh.get("b_int") == 0)


def _tear_bar(kind, precision, maxdmg=0.3):
    pos, vol = pd.lattice_block(16, 8, 8, 1e-3, jitter_amplitude=0.05, seed=7)
    mat = pd.CMaterial(maxdmg=maxdmg) if kind == "CPD" else MATS[kind]()
    b = pd.Body(mat, pos, vol)
    pd.material(b, horizon=3.015e-3, rho=7850.0, E=210e9, nu=0.25, epsilon_c=2e-3)
    x0 = pos[0]
    pd.point_set(b, "end", x0 > x0.max() - 2.5e-3)
    pd.point_set(b, "fixed", x0 < x0.min() + 2.5e-3)
    pd.velocity_bc(lambda t: 30.0, b, "end", "x")
    for d in "xyz":
        pd.velocity_bc(lambda t: 0.0, b, "fixed", d)
    bcs = [("sd", b.point_sets["end"], 0, lambda t: 30.0)]
    bcs += [("sd", b.point_sets["fixed"], d, lambda t: 0.0) for d in range(3)]
    return b, bcs


@pytest.mark.parametrize("kind", KINDS)
def test_tearing_bar_trajectory(need_gpu, oracle, kind):
    """A bar whose end is pulled at 30 m/s tears within 80 steps; for CMaterial (maxdmg 0.3)
    points along the tear are killed during the run.  Trajectory, bond states, n_active and
    damage against the oracle."""
    say("tearing bar", kind)
    b, bcs = _tear_bar(kind, "fp64")
    steps = 80
    vv = pd.VelocityVerlet(steps=steps)
    dh = pd.MI355XBodyDataHandler(b)
    dh.init_time_solver(vv)
    dh.initialize(vv)
    dh.solve(vv)
    orc = oracle_of(oracle, b, kind)
    assert vv.dt == orc.stable_dt(0.7)
    orc.force()
    orc.vv_run(steps, vv.dt, bcs)
    for f, ref, tol in (("displacement", orc.u, 1e-9), ("velocity", orc.v, 1e-8)):
        err = normwise(dh.get(f), ref)
        say(f"  {f} normwise error {err:.3e}")
        assert err < tol
    killed = int(np.sum(orc.n_active == 0))
    broken = int(np.sum(orc.active == 0))
    say(f"  oracle: {broken} inactive bonds, {killed} points without an active bond")
    assert broken > 0
    if kind == "CPD":
        assert killed > 0
    states_equal(dh, orc)


def test_tearing_bar_cpd_mixed(need_gpu, oracle):
    """The mixed-precision CPD passes with kills: the force of the last step within the
    mixed bar and the bond states equal to the FP64 oracle after a short run."""
    say("tearing bar cpd mixed")
    b, bcs = _tear_bar("CPD", "mixed")
    steps = 40
    vv = pd.VelocityVerlet(steps=steps)
    dh = pd.MI355XBodyDataHandler(b, precision="mixed")
    dh.init_time_solver(vv)
    dh.initialize(vv)
    dh.solve(vv)
    orc = oracle_of(oracle, b, "CPD")
    orc.force()
    orc.vv_run(steps, vv.dt, bcs)
    err = normwise(dh.get("displacement"), orc.u)
    _, _, _, _, act = dh.bonds()
    mism = int(np.sum(act != orc.active))
    killed = int(np.sum(orc.n_active == 0))
    say(f"  displacement normwise error {err:.3e}, broken-bond mismatches {mism}, "
        f"oracle kills {killed}")
    assert killed > 0
    assert err < 1e-5
    assert mism <= max(1, act.size // 10000)


# ------------------------------------------------------------------ Cauchy stress
@pytest.mark.parametrize("model", [pd.SaintVenantKirchhoff, pd.NeoHookePenalty,
                                   pd.LinearElastic, pd.NeoHooke])
def test_cauchy_stress_matches_oracle(need_gpu, oracle, model):
    """cauchy_stress (stress.jl:112-116, sigma = P F' / J) stored by the CPD pass of every
    point that is not killed, for the four constitutive models.  Bar: the absolute error
    over the P-wave modulus lambda + 2 mu, i.e. in units of strain, at most 1e-13: E =
    (F'F - I) / 2 of a strain of 2e-4 inherits the absolute rounding of F (~1e-15), so a
    bar relative to the stress itself would grow as 1 / strain."""
    say("cauchy stress", model.__name__)
    pos, vol = pd.lattice_block(16, 12, 8, 1e-3, jitter_amplitude=0.05, seed=5)
    b = pd.Body(pd.CMaterial(model=model()), pos, vol)
    pd.material(b, horizon=3.015e-3, rho=7850.0, E=210e9, nu=0.25, Gc=20.0)
    dh = pd.MI355XBodyDataHandler(b)
    orc = oracle_of(oracle, b, "CPD")
    X = b.xyz
    rng = np.random.default_rng(9)
    A = np.eye(3) + 2e-4 * rng.standard_normal((3, 3))
    x = X @ A.T + 1e-7 * np.sin(X * 700.0)
    dh.set("position", x)
    dh.calc_force_density(0.0, 1e-7)
    orc.force(x)
    live = orc.n_active > 0
    sg = dh.get("cauchy_stress")
    p = b.point_params[0]
    err = float(np.max(np.abs(sg[live] - orc.cauchy[live])) / (p.lam + 2.0 * p.mu))
    say(f"  cauchy error {err:.3e} strain units (normwise "
        f"{normwise(sg[live], orc.cauchy[live]):.3e}) over {int(live.sum())} live points")
    assert err <= STRESS_TOL
    assert np.abs(orc.cauchy[live]).max() > 1e6  # a real stress state, not zeros
    assert scaled_error(dh.get("b_int"), orc) <= FP64_TOL


# ------------------------------------------------------------------ reference fixture
def test_osb_interface_fixture(need_gpu, golden):
    """The reference's known answer across an OSB material interface
    (test_ordinary_state_based.jl:31-61): two points of two parameter sets, E_b = 2 E_a,
    volumes [1, 2]."""
    g = golden["osb_interface"]
    say("osb interface fixture")
    pos = np.array(g["position"])
    b = pd.Body(pd.OSBMaterial(), pos, np.array(g["volume"]))
    pd.point_set(b, "a", [0])
    pd.point_set(b, "b", [1])
    for name, E in (("a", g["E_a"]), ("b", g["E_b"])):
        pd.material(b, name, horizon=g["horizon"], rho=g["rho"], E=E, nu=g["nu"], Gc=g["Gc"])
    pd.no_failure(b)
    dh = pd.MI355XBodyDataHandler(b)
    x = b.xyz.copy()
    x[1, 0] = g["moved_x_of_point_1"]
    dh.set("position", x)
    dh.calc_force_density(0.0, 0.0)
    bi = dh.get("b_int")
    np.testing.assert_allclose(bi[0], [g["b1"], 0, 0], rtol=1e-13, atol=0)
    np.testing.assert_allclose(bi[1], [-g["b2"], 0, 0], rtol=1e-13, atol=0)
