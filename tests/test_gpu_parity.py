"""Parity of the HIP path (libpdx through its C ABI) with the CPU oracle.

Bars: bond lists bit-exact; FP64 force density within 1e-12 of the oracle, measured per
point against the sum of the absolute bond contributions to that point (the rounding scale
of a sum of ~120 nearly cancelling terms; a per-component relative error is meaningless
where the terms cancel); mixed precision within 1e-5 on the same scale, with the
broken-bond mismatch reported.  Multi-step trajectories: displacement within 1e-9 and
velocity within 1e-8 of the oracle, normwise (rounding differences of the force sums
grow over the steps), and no broken-bond mismatch."""
import sys

import numpy as np
import pytest

import peridynamics as pd
from conftest import as_oracle_params, bond_layout

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-5


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def body_block(kind, nx, ny, nz, dx=1e-3, jit=0.0, seed=0x5EED, Gc=100.0, E=70e9, nu=None,
               rho=2700.0, model=None, fail=True, zem=None):
    pos, vol = pd.lattice_block(nx, ny, nz, dx, jitter_amplitude=jit, seed=seed)
    ckw = {} if zem is None else dict(zem=zem)
    mat = {"BB": lambda: pd.BBMaterial(), "OSB": lambda: pd.OSBMaterial(),
           "CPD": lambda: pd.CMaterial(model=model, **ckw)}[kind]()
    b = pd.Body(mat, pos, vol)
    kw = dict(horizon=3.015 * dx, rho=rho, E=E)
    if kind != "BB":
        kw["nu"] = 0.25 if nu is None else nu
    if Gc is not None:
        kw["Gc"] = Gc
    pd.material(b, **kw)
    if not fail:
        pd.no_failure(b)
    return b


def oracle_for(oracle, body, kind, dh=None):
    p = body.point_params[0]
    extra = {}
    if kind == "CPD":
        names = {pd.NeoHookePenalty: "NHP", pd.LinearElastic: "LE", pd.NeoHooke: "NH",
                 pd.SaintVenantKirchhoff: "SVK"}
        wan = isinstance(body.mat.zem, pd.ZEMWan)
        extra = dict(model=names[type(body.mat.constitutive_model)],
                     Cs=0.0 if wan else body.mat.zem.Cs, maxdmg=body.mat.maxdmg,
                     zem="Wan" if wan else "Silling")
    cracks = [(body.point_sets[a], body.point_sets[b], f) for a, b, f in body.point_sets_precracks]
    return oracle.OracleSim(body.xyz, body.volume, kind, as_oracle_params(p, **extra),
                            fail_permit=body.fail_permit.astype(np.uint8), precracks=cracks)


def displacement_field(X, amp):
    L = X.max(0) - X.min(0) + 1e-30
    s = (X - X.min(0)) / L
    u = np.stack([np.sin(2.1 * s[:, 1] + 0.3), np.cos(1.7 * s[:, 2] + 0.5 * s[:, 0]),
                  np.sin(2.9 * s[:, 0] * s[:, 1])], 1)
    return amp * u


def normwise(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def scaled_error(b_gpu, orc):
    """max over points of |b_gpu - b_oracle| / (sum over the point's bond terms of |term|)."""
    scale = np.max(orc.b_abs, axis=1)
    diff = np.max(np.abs(b_gpu - orc.b_int), axis=1)
    assert np.all(diff[scale == 0] == 0)
    m = scale > 0
    return float(np.max(diff[m] / scale[m])) if m.any() else 0.0


# Cauchy stress: absolute error over the P-wave modulus lambda + 2 mu, i.e. the error in
# units of strain.  E = (F'F - I) / 2 of a small strain inherits the absolute rounding of F
# (~1e-15), so a relative bar on the stress itself would scale with 1 / strain.
STRESS_TOL = 1e-13


def stress_error(s_gpu, s_orc, p):
    return float(np.max(np.abs(s_gpu - s_orc)) / (p.lam + 2.0 * p.mu))


TRAJ_U_TOL = 1e-9
TRAJ_V_TOL = 1e-8


def bonds_equal(dh, orc):
    row_ptr, nbr, L, fp, act = dh.bonds()
    assert np.array_equal(row_ptr, orc.row_ptr)
    assert np.array_equal(nbr, orc.nbr)
    assert np.array_equal(L, orc.L)  # bit-exact bond lengths
    assert np.array_equal(fp, orc.bfp)
    return act


# ------------------------------------------------------------------ bond system
def test_tetra_bond_list(need_gpu, golden):
    g = golden["tetra_bonds"]
    say("tetra bonds")
    b = pd.Body(pd.BBMaterial(), np.array(g["position"]), np.ones(4))
    pd.material(b, horizon=g["horizon"], rho=1, E=1, Gc=1)
    dh = pd.MI355XBodyDataHandler(b)
    row_ptr, nbr, L, fp, act = dh.bonds()
    assert [[int(j), float(l)] for j, l in zip(nbr, L)] == g["bonds"]
    assert list(np.diff(row_ptr)) == g["n_neighbors"]
    assert np.all(act == 1) and np.all(fp == 1)


def test_box_bond_count(need_gpu, golden):
    g = golden["box_4x4x4_bonds"]
    say("box 4x4x4 bonds")
    pos, vol = pd.uniform_box(*g["uniform_box"])
    b = pd.Body(pd.BBMaterial(), pos, vol)
    pd.material(b, horizon=g["horizon"], rho=1, E=1, Gc=1)
    dh = pd.MI355XBodyDataHandler(b)
    assert dh.n_bonds == g["n_bonds"]
    assert np.all(dh.get("n_active_bonds") == g["n_active_per_point"])
    assert np.all(dh.get("damage") == 0)


@pytest.mark.parametrize("shape,jit", [((40, 40, 4), 0.0), ((23, 17, 11), 0.05),
                                       ((64, 48, 9), 0.05)])
def test_bond_lists_bit_exact(need_gpu, oracle, shape, jit):
    say("bond lists", shape, jit)
    b = body_block("BB", *shape, jit=jit)
    dh = pd.MI355XBodyDataHandler(b)
    orc = oracle_for(oracle, b, "BB")
    act = bonds_equal(dh, orc)
    assert np.all(act == 1)


@pytest.mark.parametrize("update_dmg", [True, False])
def test_precrack_bonds(need_gpu, oracle, update_dmg):
    say("precrack", update_dmg)
    b = body_block("BB", 30, 30, 3, jit=0.05)
    X = b.position
    pd.point_set(b, "a", (X[0] < 0) & (X[1] > 0) & (X[1] < 2.1e-3))
    pd.point_set(b, "b", (X[0] < 0) & (X[1] <= 0) & (X[1] > -2.1e-3))
    pd.precrack(b, "a", "b", update_dmg=update_dmg)
    dh = pd.MI355XBodyDataHandler(b)
    orc = oracle_for(oracle, b, "BB")
    act = bonds_equal(dh, orc)
    assert np.array_equal(act, orc.active)
    assert np.array_equal(dh.get("n_active_bonds").astype(np.int64), orc.n_active)
    np.testing.assert_array_equal(dh.get("damage"), orc.damage)
    if update_dmg:
        assert np.any(orc.damage > 0)


def test_duplicate_point(need_gpu):
    say("duplicate point")
    pos = np.array([[0.0, 1.0, 1.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    b = pd.Body(pd.BBMaterial(), pos, np.ones(3))
    pd.material(b, horizon=1.5, rho=1, E=1)
    with pytest.raises(pd.DuplicatePointError):
        pd.MI355XBodyDataHandler(b)


def test_stable_timestep_bit_exact(need_gpu, oracle):
    say("stable dt")
    b = body_block("BB", 20, 14, 6, jit=0.05)
    dh = pd.MI355XBodyDataHandler(b)
    orc = oracle_for(oracle, b, "BB")
    assert dh.calc_stable_timestep(0.7) == orc.stable_dt(0.7)


# ------------------------------------------------------------------ force density
@pytest.mark.parametrize("kind", ["BB", "OSB", "CPD"])
def test_force_density_nonuniform_volume(need_gpu, oracle, kind):
    """Point volumes varying by +-30 %: the state-based gathers must weight the term a
    neighbour j scatters into i by V_j, as the reference's b_int[j] -= p V_i of j's loop
    does (ordinary_state_based.jl:145-146, correspondence.jl:248-249)."""
    say("force nonuniform volume", kind)
    E = 70e9 if kind == "BB" else 210e9
    b = body_block(kind, 16, 14, 6, jit=0.05, E=E, Gc=None, rho=7850.0)
    rng = np.random.default_rng(3)
    b.volume *= rng.uniform(0.7, 1.3, b.n_points)
    dh = pd.MI355XBodyDataHandler(b)
    orc = oracle_for(oracle, b, kind)
    x = b.xyz + displacement_field(b.xyz, 1e-6)
    dh.set("position", x)
    dh.calc_force_density(0.0, 1e-7)
    orc.force(x)
    err = scaled_error(dh.get("b_int"), orc)
    say(f"  scaled error {err:.3e}")
    assert err <= FP64_TOL


@pytest.mark.parametrize("layout", ["lds", "lds12k", "global"])
@pytest.mark.parametrize("kind", ["BB", "OSB", "CPD"])
@pytest.mark.parametrize("amp", [2e-7, 3e-6])
def test_force_density_fp64(need_gpu, oracle, kind, amp, layout):
    """layout lds12k: a 12 KB LDS budget, so the neighbour tables are longer than what a
    block stages at once (staged in tiles, or read past the budget from global memory);
    global: the global-gather kernels."""
    say("force fp64", kind, amp, layout)
    E = 70e9 if kind == "BB" else 210e9
    # Gc 10 J/m^2 (BB, eps_c = 2.0e-4) and 20 J/m^2 (OSB, CPD, eps_c = 1.6e-4): amp 3e-6
    # breaks a tenth of the bonds, amp 2e-7 none
    b = body_block(kind, 24, 20, 8, jit=0.05, E=E, Gc=10.0 if kind == "BB" else 20.0,
                   rho=2700.0 if kind == "BB" else 7850.0)
    with bond_layout(layout):
        dh = pd.MI355XBodyDataHandler(b)
    assert dh.ctx.layout()["lds_mode"] == (layout != "global")
    orc = oracle_for(oracle, b, kind)
    X = b.xyz
    x = X + displacement_field(X, amp)
    dh.set("position", x)
    dh.calc_force_density(0.0, 1e-7)
    orc.force(x)
    bg = dh.get("b_int")
    err = scaled_error(bg, orc)
    say(f"  scaled error {err:.3e}, normwise {normwise(bg, orc.b_int):.3e}")
    assert err <= FP64_TOL
    np.testing.assert_array_equal(dh.get("n_active_bonds").astype(np.int64), orc.n_active)
    np.testing.assert_array_equal(dh.get("damage"), orc.damage)
    _, _, _, _, act = dh.bonds()
    oact = orc.active.copy()
    assert np.array_equal(act, oact)
    if amp > 1e-6:
        assert np.sum(oact == 0) > oact.size // 20  # the failure branch ran
    if kind == "CPD":
        dg = dh.get("defgrad")
        fin = np.isfinite(orc.defgrad).all(1)
        assert normwise(dg[fin], orc.defgrad[fin]) < 1e-12
        live = orc.n_active > 0
        p = b.point_params[0]
        assert stress_error(dh.get("cauchy_stress")[live], orc.cauchy[live], p) <= STRESS_TOL


@pytest.mark.parametrize("layout", ["lds", "lds12k"])
@pytest.mark.parametrize("model", [None, pd.NeoHookePenalty])
@pytest.mark.parametrize("amp", [2e-7, 3e-6])
def test_force_density_cpd_zemwan(need_gpu, oracle, amp, model, layout):
    """CMaterial(zem=ZEMWan()): the stabilisation w C_1 z with C_1 = C : K^-1 of each
    point (correspondence.jl:253-285), against the oracle at the FP64 tolerance; with
    Gc = 20 J/m^2, amp 3e-6 breaks about an eighth of the bonds."""
    say("force cpd zemwan", amp, model, layout)
    b = body_block("CPD", 24, 20, 8, jit=0.05, E=210e9, Gc=20.0, rho=7850.0,
                   model=None if model is None else model(), zem=pd.ZEMWan())
    with bond_layout(layout):
        dh = pd.MI355XBodyDataHandler(b)
    orc = oracle_for(oracle, b, "CPD")
    X = b.xyz
    x = X + displacement_field(X, amp)
    dh.set("position", x)
    dh.calc_force_density(0.0, 1e-7)
    orc.force(x)
    err = scaled_error(dh.get("b_int"), orc)
    say(f"  scaled error {err:.3e}")
    assert err <= FP64_TOL
    np.testing.assert_array_equal(dh.get("damage"), orc.damage)
    _, _, _, _, act = dh.bonds()
    assert np.array_equal(act, orc.active)
    if amp > 1e-6:
        assert np.sum(orc.active == 0) > orc.active.size // 20


def test_cpd_zemwan_mixed_context_and_global_layout(need_gpu, oracle):
    """A mixed-precision context with ZEMWan runs the FP64 passes (FP64 tolerance); the
    global-gather bond layout refuses ZEMWan with an error."""
    say("cpd zemwan mixed context / global layout")
    b = body_block("CPD", 16, 12, 8, jit=0.05, E=210e9, Gc=None, rho=7850.0, zem=pd.ZEMWan())
    dh = pd.MI355XBodyDataHandler(b, precision="mixed")
    orc = oracle_for(oracle, b, "CPD")
    x = b.xyz + displacement_field(b.xyz, 3e-6)
    dh.set("position", x)
    dh.calc_force_density(0.0, 1e-7)
    orc.force(x)
    assert scaled_error(dh.get("b_int"), orc) <= FP64_TOL
    with bond_layout("global"):
        with pytest.raises(RuntimeError, match="ZEMWan"):
            pd.MI355XBodyDataHandler(b)


def test_force_density_cpd_nhp_fixture(need_gpu, golden):
    g = golden["cpd_nhp_displaced_point"]
    say("cpd nhp fixture")
    b = pd.Body(pd.CMaterial(model=pd.NeoHookePenalty()), np.array(g["position"]),
                np.array(g["volume"]))
    pd.material(b, horizon=g["horizon"], rho=g["rho"], E=g["E"], nu=g["nu"], Gc=g["Gc"])
    pd.no_failure(b)
    dh = pd.MI355XBodyDataHandler(b)
    x = b.xyz.copy()
    x[1, 0] += g["shift_x_of_point_1"]
    dh.set("position", x)
    dh.calc_force_density(0.0, 0.0)
    bi = dh.get("b_int")
    tol = dict(rtol=0, atol=1e-11 * np.abs(bi).max())
    np.testing.assert_allclose(bi[0], g["b_int_point_0"], **tol)
    np.testing.assert_allclose(bi[1, 0], g["b_int_x_point_1"], **tol)
    np.testing.assert_allclose(bi[2, 1], g["b_int_y_point_2"], **tol)
    assert np.all(bi[4] == 0)


def test_osb_known_answer(need_gpu, golden):
    g = golden["osb_stretched_bond"]
    say("osb known answer")
    b = pd.Body(pd.OSBMaterial(), np.array(g["position"]), np.array(g["volume"]))
    pd.material(b, horizon=g["horizon"], rho=g["rho"], E=g["E"], nu=g["nu"], Gc=g["Gc"])
    pd.no_failure(b)
    dh = pd.MI355XBodyDataHandler(b)
    x = b.xyz.copy()
    x[1, 0] = g["moved_x_of_point_1"]
    dh.set("position", x)
    dh.calc_force_density(0.0, 0.0)
    bi = dh.get("b_int")
    np.testing.assert_allclose(bi[0], [g["b12"], 0, 0], rtol=1e-13, atol=0)
    np.testing.assert_allclose(bi[1], [-g["b12"], 0, 0], rtol=1e-13, atol=0)


@pytest.mark.parametrize("layout", ["lds", "lds12k", "global"])
def test_mixed_force_error(need_gpu, oracle, layout):
    say("force mixed", layout)
    b = body_block("BB", 24, 20, 8, jit=0.05)
    with bond_layout(layout):
        dh = pd.MI355XBodyDataHandler(b, precision="mixed")
    orc = oracle_for(oracle, b, "BB")
    X = b.xyz
    x = X + displacement_field(X, 2e-7)
    dh.set("position", x)
    dh.calc_force_density(0.0, 1e-7)
    orc.force(x)
    bg = dh.get("b_int")
    err = scaled_error(bg, orc)
    _, _, _, _, act = dh.bonds()
    mism = int(np.sum(act != orc.active))
    say(f"  mixed scaled error {err:.3e}, normwise {normwise(bg, orc.b_int):.3e}, "
        f"broken-bond mismatches {mism}")
    assert err <= MIXED_TOL
    assert mism <= max(1, orc.active.size // 100000)


def test_nan_position_raises(need_gpu):
    say("nan")
    b = body_block("BB", 8, 8, 4)
    dh = pd.MI355XBodyDataHandler(b)
    x = b.xyz.copy()
    x[5, 1] = np.nan
    dh.set("position", x)
    with pytest.raises(pd.NaNError):
        dh.calc_force_density(0.0, 1e-7)


# ------------------------------------------------------------------ Velocity Verlet
def _cfg1_bcs(b):
    y = b.position[1]
    pd.point_set(b, "top", y > y.max() - 2.5e-3)
    pd.point_set(b, "bottom", y < y.min() + 2.5e-3)
    pd.velocity_bc(lambda t: 0.6, b, "top", "y")
    pd.velocity_bc(lambda t: -0.6, b, "bottom", "y")
    return [("sd", b.point_sets["top"], 1, lambda t: 0.6),
            ("sd", b.point_sets["bottom"], 1, lambda t: -0.6)]


def test_vv_cfg1_fp64(need_gpu, oracle):
    """Configuration 1: 40x40x4 lattice, BB, +-0.6 m/s on the top/bottom 3 rows, 500 steps."""
    say("vv cfg1 fp64 500 steps")
    b = body_block("BB", 40, 40, 4)
    bcs = _cfg1_bcs(b)
    vv = pd.VelocityVerlet(steps=500)
    dh = pd.MI355XBodyDataHandler(b)
    dh.init_time_solver(vv)
    orc = oracle_for(oracle, b, "BB")
    assert vv.dt == orc.stable_dt(0.7)
    dh.initialize(vv)
    dh.solve(vv)
    orc.force()  # initialize!
    orc.vv_run(500, vv.dt, bcs)
    for f, ref, tol in (("displacement", orc.u, TRAJ_U_TOL), ("velocity", orc.v, TRAJ_V_TOL)):
        err = normwise(dh.get(f), ref)
        say(f"  {f} normwise error {err:.3e}")
        assert err < tol
    _, _, _, _, act = dh.bonds()
    mism = int(np.sum(act != orc.active))
    say(f"  broken bonds gpu {int(np.sum(act == 0))} oracle {int(np.sum(orc.active == 0))}, mismatches {mism}")
    assert mism == 0
    np.testing.assert_array_equal(dh.get("damage"), orc.damage)


@pytest.mark.parametrize("kind", ["OSB", "CPD"])
def test_vv_state_based_torsion(need_gpu, oracle, kind):
    """Configuration-3 style torsion on a small bar: v_y = -10 z, v_z = 10 y at one end."""
    say("vv torsion", kind)
    b = body_block(kind, 16, 8, 8, dx=1e-3, E=210e9, rho=7850.0, Gc=None, jit=0.02)
    x0 = b.position[0]
    pd.point_set(b, "end", x0 > x0.max() - 2.5e-3)
    pd.point_set(b, "fixed", x0 < x0.min() + 2.5e-3)
    pd.velocity_bc(lambda p, t: -10.0 * p[2], b, "end", "y")
    pd.velocity_bc(lambda p, t: 10.0 * p[1], b, "end", "z")
    for d in "xyz":
        pd.velocity_bc(lambda t: 0.0, b, "fixed", d)
    vv = pd.VelocityVerlet(steps=40)
    dh = pd.MI355XBodyDataHandler(b)
    dh.init_time_solver(vv)
    dh.initialize(vv)
    dh.solve(vv)
    orc = oracle_for(oracle, b, kind)
    orc.force()
    bcs = [("pd", b.point_sets["end"], 1, lambda p, t: -10.0 * p[2]),
           ("pd", b.point_sets["end"], 2, lambda p, t: 10.0 * p[1])]
    bcs += [("sd", b.point_sets["fixed"], d, lambda t: 0.0) for d in range(3)]
    orc.vv_run(40, vv.dt, bcs)
    for f, ref, tol in (("displacement", orc.u, TRAJ_U_TOL), ("velocity", orc.v, TRAJ_V_TOL)):
        err = normwise(dh.get(f), ref)
        say(f"  {f} normwise error {err:.3e}")
        assert err < tol


def test_vv_reproducible(need_gpu):
    say("vv reproducible")
    outs = []
    for _ in range(2):
        b = body_block("BB", 20, 20, 6, jit=0.05)
        _cfg1_bcs(b)
        vv = pd.VelocityVerlet(steps=50)
        dh = pd.MI355XBodyDataHandler(b)
        dh.init_time_solver(vv)
        dh.initialize(vv)
        dh.solve(vv)
        outs.append(dh.get("displacement"))
        dh.close()
    assert np.array_equal(outs[0], outs[1])


def test_vv_chunked_equals_single_run(need_gpu):
    say("vv chunked")
    res = []
    for chunk in (None, 7):
        b = body_block("BB", 20, 20, 6, jit=0.05)
        _cfg1_bcs(b)
        vv = pd.VelocityVerlet(steps=30)
        dh = pd.MI355XBodyDataHandler(b)
        dh.init_time_solver(vv)
        dh.initialize(vv)
        dh.solve(vv, max_chunk=chunk)
        res.append((dh.get("displacement"), dh.get("velocity")))
        dh.close()
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
