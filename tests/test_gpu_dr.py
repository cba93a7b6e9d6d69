"""DynamicRelaxation on the device against the reference's known answers
(test/unit/time_solvers/test_dynamic_relaxation.jl) and the oracle's trajectory."""
import sys

import numpy as np
import pytest

import peridynamics as pd
from conftest import as_oracle_params

pytestmark = pytest.mark.gpu

TETRA = np.array([[0.0, 1.0, 0.0, 0.0], [0.0, 0.0, 0.0, 1.0], [0.0, 0.0, 1.0, 0.0]])


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("E,expected", [(210e3, 4.2e6), (200e3, 4.0e6)])
def test_density_matrix_known_answer(need_gpu, E, expected):
    """_init_density_matrix (test_dynamic_relaxation.jl:24-43): tetra, horizon 1.5, three
    bonds per point, Lambda = 4.2e6 for E = 210e3 and 4.0e6 for E = 200e3."""
    say("dr density", E)
    b = pd.Body(pd.BBMaterial(), TETRA, np.ones(4))
    pd.material(b, horizon=1.5, rho=8e-6, E=E)
    dh = pd.MI355XBodyDataHandler(b)
    dh.ctx.dr_init(1.0, 1.0)
    np.testing.assert_allclose(dh.get("density_matrix"), expected, rtol=1e-12)


def test_two_steps_known_answer(need_gpu):
    """relaxation_timestep! twice (test_dynamic_relaxation.jl:58-111) with one parameter set:
    horizon 3, E = 200e3 (Lambda = 1e6 for every point), b_ext = 1e5 x on points 1 and 2 ->
    after two steps point 2 sits at x = 1 + dt * dt/2 * 1e5 / 1e6 = 1.05."""
    say("dr two steps")
    b = pd.Body(pd.BBMaterial(), TETRA, np.ones(4))
    pd.material(b, horizon=3.0, rho=7.8e-6, E=200e3)
    pd.point_set(b, "a", [0, 1])
    pd.forcedensity_bc(lambda p, t: 1e5 * p[0], b, "a", "x")
    dr = pd.DynamicRelaxation(steps=2)
    dh = pd.submit(pd.Job(b, dr, fields=()))
    np.testing.assert_allclose(dh.get("density_matrix"), 1e6, rtol=1e-12)
    assert dh.get("position")[1, 0] == pytest.approx(1.05, rel=1e-12)


def test_damping_clamped(need_gpu):
    """calc_damping (test_dynamic_relaxation.jl:120-143) on the line10 fixture."""
    say("dr damping")
    pos = np.zeros((3, 10))
    pos[0] = np.arange(10.0)
    b = pd.Body(pd.BBMaterial(), pos, np.ones(10))
    pd.material(b, horizon=1.5, rho=1.0, E=1.0, nu=0.25)
    dh = pd.MI355XBodyDataHandler(b)
    dh.ctx.dr_init(1.0, 1.0)
    dm = dh.get("density_matrix")
    assert np.all(dm > 0)
    for f, v in (("displacement", 1.0), ("velocity_half", 1.0), ("velocity_half_old", 1.0),
                 ("b_int_old", 0.0), ("b_int", -1e6)):
        dh.set(f, np.full((10, 3), v))
    assert dh.calc_damping(1.0) == 1.9
    dh.set("b_int", np.full((10, 3), -1e-6))
    cn1 = sum(1e-6 / d for d in np.repeat(dm, 3))  # all 30 dofs contribute
    assert dh.calc_damping(1.0) == pytest.approx(2 * np.sqrt(cn1 / 30), rel=1e-14)
    dh.set("displacement", np.zeros((10, 3)))
    assert dh.calc_damping(1.0) == 0.0


def test_dr_trajectory_matches_oracle(need_gpu, oracle):
    """A 12 x 12 x 4 block pulled at 0.01 per step on its top rows, bottom fixed, 150 steps."""
    say("dr trajectory")
    pos, vol = pd.lattice_block(12, 12, 4, 1e-3, jitter_amplitude=0.05, seed=7)
    b = pd.Body(pd.BBMaterial(), pos, vol)
    pd.material(b, horizon=3.015e-3, rho=2700.0, E=70e9)
    y = pos[1]
    pd.point_set(b, "top", y > y.max() - 1.5e-3)
    pd.point_set(b, "bottom", y < y.min() + 1.5e-3)
    pd.velocity_bc(lambda t: 1e-6, b, "top", "y")
    for d in "xyz":
        pd.velocity_bc(lambda t: 0.0, b, "bottom", d)
    dr = pd.DynamicRelaxation(steps=150, stepsize=1.0)
    dh = pd.MI355XBodyDataHandler(b)
    dh.calc_force_density(0.0, 1.0)
    dh.solve_dr(dr)
    orc = oracle.OracleSim(b.xyz, b.volume, "BB", as_oracle_params(b.point_params[0]),
                           fail_permit=b.fail_permit.astype(np.uint8))
    bcs = [("sd", b.point_sets["top"], 1, lambda t: 1e-6)] + \
          [("sd", b.point_sets["bottom"], d, lambda t: 0.0) for d in range(3)]
    orc.dr_run(150, 1.0, 1.0, bcs)
    for f, ref in (("displacement", orc.u), ("velocity_half", orc.vh)):
        g = dh.get(f)
        err = np.max(np.abs(g - ref)) / np.max(np.abs(ref))
        say(f"  {f} normwise error {err:.2e}")
        assert err < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kind", ["BB", "OSB", "CPD", "CPD-ZEMWan"])
def test_dr_materials(need_gpu, oracle, kind, precision):
    """Dynamic relaxation of BB, OSB and CPD bodies (calc_density_matrix! for any bond
    system, dynamic_relaxation.jl:153-163): the block of the trajectory test against the
    FP64 oracle.  The adaptive damping divides by differences of consecutive force
    densities, which amplifies rounding differences from step to step; the state-based
    forces, sums of terms that cancel more, are compared over 20 steps, BB over 150."""
    say("dr", kind, precision)
    n_steps = 150 if kind == "BB" else 20
    pos, vol = pd.lattice_block(12, 12, 4, 1e-3, jitter_amplitude=0.05, seed=7)
    wan = kind == "CPD-ZEMWan"
    if wan:
        kind = "CPD"
    mat = {"BB": pd.BBMaterial, "OSB": pd.OSBMaterial,
           "CPD": lambda: pd.CMaterial(zem=pd.ZEMWan()) if wan else pd.CMaterial()}[kind]()
    b = pd.Body(mat, pos, vol)
    kw = dict(horizon=3.015e-3, rho=2700.0, E=70e9)
    if kind != "BB":
        kw["nu"] = 0.25
    pd.material(b, **kw)
    y = pos[1]
    pd.point_set(b, "top", y > y.max() - 1.5e-3)
    pd.point_set(b, "bottom", y < y.min() + 1.5e-3)
    pd.velocity_bc(lambda t: 1e-6, b, "top", "y")
    for d in "xyz":
        pd.velocity_bc(lambda t: 0.0, b, "bottom", d)
    dr = pd.DynamicRelaxation(steps=n_steps, stepsize=1.0)
    dh = pd.MI355XBodyDataHandler(b, precision=precision)
    dh.calc_force_density(0.0, 1.0)
    dh.solve_dr(dr)
    extra = {}
    if kind == "CPD":
        extra = dict(model="SVK", Cs=0.0 if wan else b.mat.zem.Cs, maxdmg=b.mat.maxdmg,
                     zem="Wan" if wan else "Silling")
    orc = oracle.OracleSim(b.xyz, b.volume, kind, as_oracle_params(b.point_params[0], **extra),
                           fail_permit=b.fail_permit.astype(np.uint8))
    bcs = [("sd", b.point_sets["top"], 1, lambda t: 1e-6)] + \
          [("sd", b.point_sets["bottom"], d, lambda t: 0.0) for d in range(3)]
    orc.dr_run(n_steps, 1.0, 1.0, bcs)
    g = dh.get("displacement")
    err = np.max(np.abs(g - orc.u)) / np.max(np.abs(orc.u))
    say(f"  displacement normwise error {err:.2e}")
    assert err < (1e-9 if precision == "fp64" else 1e-5)
    dh.close()
