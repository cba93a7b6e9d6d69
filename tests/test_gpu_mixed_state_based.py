"""Mixed-precision OSB and CPD steps (FP32 bond arithmetic from FP64 differences, FP64
sums, FP64 per-point state and Verlet tail) against the FP64 oracle of
ordinary_state_based.jl:107-204 and correspondence.jl:193-251.

Bars: force density within 1e-5 of the per-point scale (the sum of the absolute bond
contributions to the point), bond states identical to the oracle's (stretches near eps_c
are decided in FP64), and a torsion trajectory that stays within 1e-5 (normwise) of the
FP64 oracle.  Each force case also runs with a 12 KB LDS budget, where most neighbour-table
entries are read from global memory instead of LDS."""
import sys

import numpy as np
import pytest

import peridynamics as pd
from conftest import lds_budget
from test_gpu_parity import body_block, displacement_field, normwise, oracle_for, scaled_error

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-5


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("budget", [0, 12288])
@pytest.mark.parametrize("amp", [2e-7, 3e-6])
def test_mixed_osb_force(need_gpu, oracle, amp, budget):
    say("mixed osb force", amp, budget)
    b = body_block("OSB", 24, 20, 8, jit=0.05, E=210e9, rho=7850.0, Gc=20.0)
    with lds_budget(budget):
        dh = pd.MI355XBodyDataHandler(b, precision="mixed")
    orc = oracle_for(oracle, b, "OSB")
    x = b.xyz + displacement_field(b.xyz, amp)
    dh.set("position", x)
    dh.calc_force_density(0.0, 1e-7)
    orc.force(x)
    bg = dh.get("b_int")
    err = scaled_error(bg, orc)
    _, _, _, _, act = dh.bonds()
    broken = int(np.sum(orc.active == 0))
    mism = int(np.sum(act != orc.active))
    say(f"  scaled error {err:.3e}, normwise {normwise(bg, orc.b_int):.3e}, "
        f"broken {broken}, mismatches {mism}")
    assert err <= MIXED_TOL
    assert mism == 0
    if amp > 1e-6:
        assert broken > 0  # the failure path is exercised
    dh.close()


def test_mixed_osb_force_with_correction(need_gpu, oracle):
    say("mixed osb force, surface correction")
    pos, vol = pd.lattice_block(16, 14, 6, 1e-3, jitter_amplitude=0.05, seed=5)
    b = pd.Body(pd.OSBMaterial(correction=pd.EnergySurfaceCorrection), pos, vol)
    pd.material(b, horizon=3.015e-3, rho=7850.0, E=210e9, nu=0.25)
    dh = pd.MI355XBodyDataHandler(b, precision="mixed")
    p = b.point_params[0]
    from conftest import as_oracle_params
    orc = oracle.OracleSim(b.xyz, b.volume, "OSB", as_oracle_params(p),
                           fail_permit=b.fail_permit.astype(np.uint8), correction="energy")
    x = b.xyz + 2e-7 * np.random.default_rng(6).standard_normal(b.xyz.shape)
    dh.set("position", x)
    dh.calc_force_density(0.0, 1e-7)
    orc.force(x)
    err = scaled_error(dh.get("b_int"), orc)
    say(f"  scaled error {err:.3e}")
    assert err <= MIXED_TOL
    dh.close()


def test_mixed_osb_torsion_trajectory(need_gpu, oracle):
    """The torsion bar of the state-based trajectory test, 40 steps, mixed precision."""
    say("mixed osb torsion")
    b = body_block("OSB", 16, 8, 8, dx=1e-3, E=210e9, rho=7850.0, Gc=None, jit=0.02)
    x0 = b.position[0]
    pd.point_set(b, "end", x0 > x0.max() - 2.5e-3)
    pd.point_set(b, "fixed", x0 < x0.min() + 2.5e-3)
    pd.velocity_bc(lambda p, t: -10.0 * p[2], b, "end", "y")
    pd.velocity_bc(lambda p, t: 10.0 * p[1], b, "end", "z")
    for d in "xyz":
        pd.velocity_bc(lambda t: 0.0, b, "fixed", d)
    vv = pd.VelocityVerlet(steps=40)
    dh = pd.MI355XBodyDataHandler(b, precision="mixed")
    dh.init_time_solver(vv)
    dh.initialize(vv)
    dh.solve(vv)
    orc = oracle_for(oracle, b, "OSB")
    orc.force()
    bcs = [("pd", b.point_sets["end"], 1, lambda p, t: -10.0 * p[2]),
           ("pd", b.point_sets["end"], 2, lambda p, t: 10.0 * p[1])]
    bcs += [("sd", b.point_sets["fixed"], d, lambda t: 0.0) for d in range(3)]
    orc.vv_run(40, vv.dt, bcs)
    for f, ref in (("displacement", orc.u), ("velocity", orc.v)):
        err = normwise(dh.get(f), ref)
        say(f"  {f} normwise error {err:.3e}")
        assert err < 1e-5
    dh.close()


CPD_MODELS = {"SVK": pd.SaintVenantKirchhoff, "NHP": pd.NeoHookePenalty,
              "LE": pd.LinearElastic, "NH": pd.NeoHooke}


@pytest.mark.parametrize("model", ["SVK", "NHP", "LE", "NH"])
@pytest.mark.parametrize("budget", [0, 12288])
@pytest.mark.parametrize("amp", [2e-7, 3e-6])
def test_mixed_cpd_force(need_gpu, oracle, amp, budget, model):
    """Mixed CPD (correspondence.jl:193-251): force density, bond states and the
    deformation gradient against the FP64 oracle, with bond failure at the larger
    amplitude."""
    say("mixed cpd force", amp, budget, model)
    b = body_block("CPD", 24, 20, 8, jit=0.05, E=210e9, rho=7850.0, Gc=20.0,
                   model=CPD_MODELS[model]())
    with lds_budget(budget):
        dh = pd.MI355XBodyDataHandler(b, precision="mixed")
    orc = oracle_for(oracle, b, "CPD")
    x = b.xyz + displacement_field(b.xyz, amp)
    dh.set("position", x)
    dh.calc_force_density(0.0, 1e-7)
    orc.force(x)
    bg = dh.get("b_int")
    err = scaled_error(bg, orc)
    _, _, _, _, act = dh.bonds()
    broken = int(np.sum(orc.active == 0))
    mism = int(np.sum(act != orc.active))
    dg = dh.get("defgrad")
    fin = np.isfinite(orc.defgrad).all(1)
    dgerr = normwise(dg[fin], orc.defgrad[fin])
    say(f"  scaled error {err:.3e}, normwise {normwise(bg, orc.b_int):.3e}, "
        f"broken {broken}, mismatches {mism}, defgrad {dgerr:.3e}")
    assert err <= MIXED_TOL
    assert mism == 0
    assert dgerr < 1e-8
    if amp > 1e-6:
        assert broken > 0
    dh.close()


def test_mixed_cpd_torsion_trajectory(need_gpu, oracle):
    """The torsion bar with CMaterial (Saint Venant-Kirchhoff, ZEMSilling), 40 steps."""
    say("mixed cpd torsion")
    b = body_block("CPD", 16, 8, 8, dx=1e-3, E=210e9, rho=7850.0, Gc=None, jit=0.02,
                   model=pd.SaintVenantKirchhoff())
    x0 = b.position[0]
    pd.point_set(b, "end", x0 > x0.max() - 2.5e-3)
    pd.point_set(b, "fixed", x0 < x0.min() + 2.5e-3)
    pd.velocity_bc(lambda p, t: -10.0 * p[2], b, "end", "y")
    pd.velocity_bc(lambda p, t: 10.0 * p[1], b, "end", "z")
    for d in "xyz":
        pd.velocity_bc(lambda t: 0.0, b, "fixed", d)
    vv = pd.VelocityVerlet(steps=40)
    dh = pd.MI355XBodyDataHandler(b, precision="mixed")
    dh.init_time_solver(vv)
    dh.initialize(vv)
    dh.solve(vv)
    orc = oracle_for(oracle, b, "CPD")
    orc.force()
    bcs = [("pd", b.point_sets["end"], 1, lambda p, t: -10.0 * p[2]),
           ("pd", b.point_sets["end"], 2, lambda p, t: 10.0 * p[1])]
    bcs += [("sd", b.point_sets["fixed"], d, lambda t: 0.0) for d in range(3)]
    orc.vv_run(40, vv.dt, bcs)
    for f, ref in (("displacement", orc.u), ("velocity", orc.v)):
        err = normwise(dh.get(f), ref)
        say(f"  {f} normwise error {err:.3e}")
        assert err < 1e-5
    dh.close()


F_AFFINE = np.array([[1.1, 0.01, 0.0], [0.0, 1.2, 0.02], [0.03, 0.0, 1.3]])


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-6)])
@pytest.mark.parametrize("model", ["SVK", "NHP", "LE", "NH"])
def test_cpd_defgrad_exact_for_affine_deformation(need_gpu, model, precision, tol):
    """test/materials/test_uniform_deformation.jl:5-41 on the device: under x = F X every
    point's deformation gradient, surface points included, equals F (mixed precision forms
    F = I + H from FP32 displacement terms: H to FP32 rounding)."""
    say("cpd affine defgrad", model, precision)
    n = 5
    dx = 1.0 / n
    pos, vol = pd.uniform_box(1.0, 1.0, 1.0, dx)
    b = pd.Body(pd.CMaterial(model=CPD_MODELS[model]()), pos, vol)
    pd.material(b, horizon=3.1 * dx, rho=7850.0, E=210e9, nu=0.25)
    pd.no_failure(b)
    dh = pd.MI355XBodyDataHandler(b, precision=precision)
    X = b.xyz
    dh.set("position", X @ F_AFFINE.T)
    dh.calc_force_density(0.0, 1e-7)
    err = np.max(np.abs(dh.get("defgrad") - F_AFFINE.T.reshape(-1)))
    say(f"  max |F - F_a| {err:.2e}")
    assert err < tol
    dh.close()


@pytest.mark.parametrize("model", ["LE", "NH"])
def test_cpd_models_force_fp64(need_gpu, oracle, model):
    """LinearElastic and NeoHooke (constitutive_models.jl:81-97, 167-175) in the FP64 CPD
    step against the oracle, to the FP64 bar."""
    say("cpd model fp64", model)
    b = body_block("CPD", 24, 20, 8, jit=0.05, E=210e9, rho=7850.0, Gc=None,
                   model=CPD_MODELS[model]())
    dh = pd.MI355XBodyDataHandler(b)
    orc = oracle_for(oracle, b, "CPD")
    x = b.xyz + displacement_field(b.xyz, 3e-6)
    dh.set("position", x)
    dh.calc_force_density(0.0, 1e-7)
    orc.force(x)
    err = scaled_error(dh.get("b_int"), orc)
    say(f"  scaled error {err:.3e}")
    assert err <= 1e-12
    dh.close()
