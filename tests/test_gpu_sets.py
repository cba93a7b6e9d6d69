"""Bodies with several parameter sets on the HIP path, against the CPU oracle.

material!(body, set_name; ...) per point set gives each point its own parameters; a bond
takes the mean of its two ends' bc (BB) or G and K (OSB), a CMaterial point the parameters
of its own set, a bond fails against its owner's eps_c, and a point integrates with its own
rho (bond_based.jl:163-180, ordinary_state_based.jl:151-174, correspondence.jl:183-191,
bond_system.jl:309-326, velocity_verlet.jl:262-275).
Bars as the single-set parity tests: FP64 force within 1e-12 of the oracle on the per-point
rounding scale, bond states bit-exact, trajectories within 1e-9 (displacement) and 1e-8
(velocity) normwise, mixed precision within 1e-5."""
import sys

import numpy as np
import pytest

import peridynamics as pd
from conftest import as_oracle_params, lds_budget

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-5
SETS = {  # E, eps_c (via Gc), rho of three materials
    "a": (210e9, 3e-4, 7850.0),
    "b": (70e9, 1e-3, 2700.0),
    "c": (140e9, 6e-4, 4500.0),
}


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def sets_body(kind, nx, ny, nz, dx=1e-3, jit=0.05, seed=11, fail=True, vol_jit=0.3):
    pos, vol = pd.lattice_block(nx, ny, nz, dx, jitter_amplitude=jit, seed=seed)
    vol = vol * np.random.default_rng(seed).uniform(1 - vol_jit, 1 + vol_jit, vol.shape[0])
    b = pd.Body({"BB": pd.BBMaterial, "OSB": pd.OSBMaterial, "CPD": pd.CMaterial}[kind](), pos, vol)
    x, y = pos[0], pos[1]
    pd.point_set(b, "a", (x <= np.median(x)) & (y <= np.median(y)))
    pd.point_set(b, "b", (x > np.median(x)) & (y <= np.median(y)))
    pd.point_set(b, "c", y > np.median(y))
    delta = 3.015 * dx
    for name, (E, epsc, rho) in SETS.items():
        kw = dict(horizon=delta, rho=rho, E=E)
        if kind != "BB":
            kw["nu"] = 0.25
        if fail:
            kw["epsilon_c"] = epsc
        pd.material(b, name, **kw)
    return b


def oracle_of(oracle, b, kind):
    extra = {}
    if kind == "CPD":
        extra = dict(model="SVK", Cs=b.mat.zem.Cs, maxdmg=b.mat.maxdmg)
    sets = [as_oracle_params(p, **extra) for p in b.point_params]
    return oracle.OracleSim(b.xyz, b.volume, kind, sets,
                            fail_permit=b.fail_permit.astype(np.uint8),
                            point_set=b.params_map - 1)


def scaled_error(b_gpu, orc):
    scale = np.max(orc.b_abs, axis=1)
    diff = np.max(np.abs(b_gpu - orc.b_int), axis=1)
    assert np.all(diff[scale == 0] == 0)
    m = scale > 0
    return float(np.max(diff[m] / scale[m])) if m.any() else 0.0


def normwise(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def deform(X, amp, seed=2):
    return X + amp * np.random.default_rng(seed).standard_normal(X.shape)


@pytest.mark.parametrize("kind", ["BB", "OSB"])
def test_sets_two_point_closed_form(need_gpu, kind):
    """Two points with different E and volumes [1, 2]: the force densities are closed forms
    in the means of the two sets' constants."""
    say("sets two points", kind)
    X = np.array([[0.0, 1.0], [0.0, 0.0], [0.0, 0.0]])
    b = pd.Body(pd.BBMaterial() if kind == "BB" else pd.OSBMaterial(), X, np.array([1.0, 2.0]))
    pd.point_set(b, "p1", [0])
    pd.point_set(b, "p2", [1])
    extra = {} if kind == "BB" else {"nu": 0.25}
    pd.material(b, "p1", horizon=1.5, rho=7850.0, E=105e9, **extra)
    pd.material(b, "p2", horizon=1.5, rho=7850.0, E=70e9, **extra)
    dh = pd.MI355XBodyDataHandler(b)
    assert dh.n_sets == 2
    x = b.xyz.copy()
    x[1, 0] += 0.0015
    e = x[1, 0] - 1.0
    dh.set("position", x)
    dh.calc_force_density(0.0, 0.0)
    bi = dh.get("b_int")
    pa, pb = b.point_params
    V1, V2 = 1.0, 2.0
    if kind == "BB":
        bc = (pa.bc + pb.bc) / 2
        b1, b2 = bc * e * V2, -bc * e * V1
    else:
        w, Gs, Ks = 1.5, pa.G + pb.G, pa.K + pb.K

        def bond(wvol, dil):
            c1 = 15.0 * Gs / (2 * wvol)
            return w * (dil * (3.0 * Ks / (2 * wvol) - c1 / 3.0) + c1 * e)

        p12 = bond(w * V2, 3.0 * e)
        p21 = bond(w * V1, 3.0 * e)
        b1, b2 = (p12 + p21) * V2, -(p12 + p21) * V1
    np.testing.assert_allclose(bi[0], [b1, 0, 0], rtol=1e-13, atol=0)
    np.testing.assert_allclose(bi[1], [b2, 0, 0], rtol=1e-13, atol=0)
    dh.close()


@pytest.mark.parametrize("budget", [0, 12288])
@pytest.mark.parametrize("kind", ["BB", "OSB", "CPD"])
def test_sets_force_fp64(need_gpu, oracle, kind, budget):
    say("sets force fp64", kind, budget)
    b = sets_body(kind, 16, 14, 6)
    with lds_budget(budget):
        dh = pd.MI355XBodyDataHandler(b)
    assert dh.n_sets == 3
    orc = oracle_of(oracle, b, kind)
    x = deform(b.xyz, 5e-7)
    for it in range(2):  # the second call sees the bonds the first one broke
        dh.set("position", x)
        dh.calc_force_density(0.0, 1e-7)
        orc.force(x)
        err = scaled_error(dh.get("b_int"), orc)
        _, _, _, _, act = dh.bonds()
        say(f"  call {it}: scaled error {err:.3e}, broken {int(np.sum(orc.active == 0))}, "
            f"mismatches {int(np.sum(act != orc.active))}")
        assert err <= FP64_TOL
        assert np.array_equal(act, orc.active)
        np.testing.assert_array_equal(dh.get("damage"), orc.damage)
        x = deform(x, 2e-7, seed=3 + it)
    assert np.any(orc.active == 0)
    dh.close()


@pytest.mark.parametrize("kind", ["BB", "OSB", "CPD"])
def test_sets_equal_match_single_set(need_gpu, kind):
    """Three sets with equal parameters give the single-set result (the means of equal
    numbers are those numbers)."""
    say("sets equal vs single", kind)
    res = []
    for multi in (False, True):
        pos, vol = pd.lattice_block(14, 12, 6, 1e-3, jitter_amplitude=0.05, seed=4)
        b = pd.Body({"BB": pd.BBMaterial, "OSB": pd.OSBMaterial, "CPD": pd.CMaterial}[kind](), pos, vol)
        kw = dict(horizon=3.015e-3, rho=7850.0, E=210e9, epsilon_c=3e-4)
        if kind != "BB":
            kw["nu"] = 0.25
        if multi:
            x = pos[0]
            pd.point_set(b, "l", x <= np.median(x))
            pd.point_set(b, "r", x > np.median(x))
            pd.material(b, "l", **kw)
            pd.material(b, "r", **dict(kw, rho=7851.0))  # a distinct set, same force constants
        else:
            pd.material(b, **kw)
        dh = pd.MI355XBodyDataHandler(b)
        assert dh.n_sets == (2 if multi else 1)
        x = deform(b.xyz, 4e-7)
        dh.set("position", x)
        dh.calc_force_density(0.0, 1e-7)
        res.append((dh.get("b_int"), dh.bonds()[4]))
        dh.close()
    assert np.array_equal(res[0][1], res[1][1])
    assert normwise(res[1][0], res[0][0]) <= 1e-15


def test_sets_mixed_force(need_gpu, oracle):
    say("sets force mixed")
    b = sets_body("BB", 16, 14, 6, fail=False)
    dh = pd.MI355XBodyDataHandler(b, precision="mixed")
    orc = oracle_of(oracle, b, "BB")
    x = deform(b.xyz, 2e-7)
    dh.set("position", x)
    dh.calc_force_density(0.0, 1e-7)
    orc.force(x)
    err = scaled_error(dh.get("b_int"), orc)
    say(f"  scaled error {err:.3e}")
    assert err <= MIXED_TOL
    dh.close()


def test_sets_stable_timestep(need_gpu, oracle):
    say("sets stable dt")
    b = sets_body("BB", 14, 12, 6)
    dh = pd.MI355XBodyDataHandler(b)
    orc = oracle_of(oracle, b, "BB")
    assert dh.calc_stable_timestep(0.7) == orc.stable_dt(0.7)
    dh.close()


@pytest.mark.parametrize("kind", ["BB", "OSB", "CPD"])
def test_sets_vv_trajectory(need_gpu, oracle, kind):
    """Tension across the interfaces: +-1 m/s on the top and bottom rows, 60 steps, per-point
    rho in the update, bonds failing against each owner's eps_c."""
    say("sets vv", kind)
    b = sets_body(kind, 16, 16, 5)
    y = b.position[1]
    pd.point_set(b, "top", y > y.max() - 2.5e-3)
    pd.point_set(b, "bottom", y < y.min() + 2.5e-3)
    pd.velocity_bc(lambda t: 1.0, b, "top", "y")
    pd.velocity_bc(lambda t: -1.0, b, "bottom", "y")
    vv = pd.VelocityVerlet(steps=60)
    dh = pd.MI355XBodyDataHandler(b)
    dh.init_time_solver(vv)
    orc = oracle_of(oracle, b, kind)
    assert vv.dt == orc.stable_dt(0.7)
    dh.initialize(vv)
    dh.solve(vv)
    orc.force()
    bcs = [("sd", b.point_sets["top"], 1, lambda t: 1.0),
           ("sd", b.point_sets["bottom"], 1, lambda t: -1.0)]
    orc.vv_run(60, vv.dt, bcs)
    for f, ref, tol in (("displacement", orc.u, 1e-9), ("velocity", orc.v, 1e-8)):
        err = normwise(dh.get(f), ref)
        say(f"  {f} normwise error {err:.3e}")
        assert err < tol
    _, _, _, _, act = dh.bonds()
    say(f"  broken bonds oracle {int(np.sum(orc.active == 0))}, "
        f"mismatches {int(np.sum(act != orc.active))}")
    assert np.array_equal(act, orc.active)
    dh.close()


def test_sets_dynamic_relaxation(need_gpu, oracle):
    say("sets dr")
    b = sets_body("BB", 12, 12, 4, fail=False, vol_jit=0.0)
    y = b.position[1]
    pd.point_set(b, "top", y > y.max() - 1.5e-3)
    pd.point_set(b, "bottom", y < y.min() + 1.5e-3)
    pd.velocity_bc(lambda t: 1e-6, b, "top", "y")
    for d in "xyz":
        pd.velocity_bc(lambda t: 0.0, b, "bottom", d)
    dr = pd.DynamicRelaxation(steps=100, stepsize=1.0)
    dh = pd.MI355XBodyDataHandler(b)
    dh.calc_force_density(0.0, 1.0)
    dh.solve_dr(dr)
    orc = oracle_of(oracle, b, "BB")
    bcs = [("sd", b.point_sets["top"], 1, lambda t: 1e-6)] + \
          [("sd", b.point_sets["bottom"], d, lambda t: 0.0) for d in range(3)]
    st = orc.dr_run(100, 1.0, 1.0, bcs)
    np.testing.assert_array_equal(np.ravel(dh.get("density_matrix")), st["density"])
    for f, ref in (("displacement", orc.u), ("velocity_half", orc.vh)):
        err = normwise(dh.get(f), ref)
        say(f"  {f} normwise error {err:.2e}")
        assert err < 1e-9
    dh.close()


def test_sets_cpd_mixed_context_runs_fp64(need_gpu, oracle):
    """A mixed-precision CMaterial body with several sets runs the FP64 passes: its force
    equals the oracle's to the FP64 bar."""
    say("sets cpd mixed context")
    b = sets_body("CPD", 16, 14, 6, fail=False)
    dh = pd.MI355XBodyDataHandler(b, precision="mixed")
    orc = oracle_of(oracle, b, "CPD")
    x = deform(b.xyz, 2e-7)
    dh.set("position", x)
    dh.calc_force_density(0.0, 1e-7)
    orc.force(x)
    err = scaled_error(dh.get("b_int"), orc)
    say(f"  scaled error {err:.3e}")
    assert err <= FP64_TOL
    dh.close()