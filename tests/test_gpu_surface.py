"""EnergySurfaceCorrection on the HIP path against the CPU oracle and the reference's
fixtures (bond_system_corrections.jl; test_bond_based.jl:181-211,
test_bond_system_corrections.jl).

The device sums a point's bond energies in its own row order and takes atan/acos/cos/sin
from the device math library, so mfactor and scfactor agree with the oracle to a few
rounding errors (bar 1e-13 relative) rather than bit for bit; the force densities are held
to the FP64 parity bar (1e-12 of the per-point bond scale), trajectories to 1e-9 / 1e-8."""
import math
import sys

import numpy as np
import pytest

import peridynamics as pd
from conftest import as_oracle_params

pytestmark = pytest.mark.gpu

ESC_TOL = 1e-13
FP64_TOL = 1e-12


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def normwise(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def scaled_error(b_gpu, orc):
    scale = np.max(orc.b_abs, axis=1)
    diff = np.max(np.abs(b_gpu - orc.b_int), axis=1)
    assert np.all(diff[scale == 0] == 0)
    m = scale > 0
    return float(np.max(diff[m] / scale[m])) if m.any() else 0.0


def mat_of(kind):
    c = pd.EnergySurfaceCorrection
    return pd.BBMaterial(correction=c) if kind == "BB" else pd.OSBMaterial(correction=c)


def oracle_of(oracle, b, kind):
    if len(b.point_params) > 1:
        sets = [as_oracle_params(p) for p in b.point_params]
        return oracle.OracleSim(b.xyz, b.volume, kind, sets,
                                fail_permit=b.fail_permit.astype(np.uint8),
                                point_set=b.params_map - 1, correction="energy")
    return oracle.OracleSim(b.xyz, b.volume, kind, as_oracle_params(b.point_params[0]),
                            fail_permit=b.fail_permit.astype(np.uint8), correction="energy")


def factors_match(dh, orc):
    sc, mf = dh.surface_correction()
    fin = np.isfinite(orc.mfactor)
    assert np.array_equal(np.isfinite(mf), fin)
    em = float(np.max(np.abs(mf[fin] - orc.mfactor[fin]) / np.abs(orc.mfactor[fin])))
    es = float(np.max(np.abs(sc - orc.scf) / np.abs(orc.scf)))
    say(f"  mfactor rel {em:.2e}, scfactor rel {es:.2e}")
    assert em <= ESC_TOL and es <= ESC_TOL
    return sc, mf


def test_two_point_fixture(need_gpu, golden):
    g = golden["bb_surface_correction_two_points"]
    say("esc two points")
    b = pd.Body(mat_of("BB"), np.array(g["position"]), np.array(g["volume"]))
    pd.material(b, horizon=g["horizon"], rho=g["rho"], E=g["E"])
    pd.no_failure(b)
    dh = pd.MI355XBodyDataHandler(b)
    sc, mf = dh.surface_correction()
    assert np.allclose(sc, g["scfactor"], rtol=math.sqrt(np.finfo(float).eps), atol=0)
    assert np.allclose(sc, 1.0125 * math.pi, rtol=4e-16, atol=0)  # closed form
    x = b.xyz.copy()
    x[1, 0] = g["moved_x_of_point_1"]
    dh.set("position", x)
    dh.calc_force_density(0.0, 0.0)
    bi = dh.get("b_int")
    E, d = g["E"], g["horizon"]
    b12 = g["scfactor"] * 18 * E / (3 * (1 - 2 * 0.25)) / (math.pi * d ** 4) * 1.0015 * 0.0015 / 1.0015
    np.testing.assert_allclose(bi[:, 0], [b12, -b12], rtol=1e-9)
    dh.close()


@pytest.mark.parametrize("kind", ["BB", "OSB"])
def test_box_factors_and_force(need_gpu, oracle, golden, kind):
    g = golden["surface_correction_box"]
    say("esc box", kind)
    pos, vol = pd.uniform_box(*g["uniform_box"])
    b = pd.Body(mat_of(kind), pos, vol)
    kw = dict(horizon=g["horizon_over_dx"] * g["uniform_box"][3], rho=g["rho"], E=g["E"])
    if kind != "BB":
        kw["nu"] = g["nu"]
    pd.material(b, **kw)
    dh = pd.MI355XBodyDataHandler(b)
    orc = oracle_of(oracle, b, kind)
    sc, mf = factors_match(dh, orc)
    r = g[kind]
    for d in range(3):
        assert abs(mf[:, d].min() - r["mfactor_min"][0]) <= r["mfactor_min"][1]
        assert abs(mf[:, d].max() - r["mfactor_max"][0]) <= r["mfactor_max"][1]
    assert abs(sc.min() - r["scfactor_min"][0]) <= r["scfactor_min"][1]
    assert abs(sc.max() - r["scfactor_max"][0]) <= r["scfactor_max"][1]
    x = b.xyz + 1e-4 * np.random.default_rng(3).standard_normal(b.xyz.shape)
    dh.set("position", x)
    dh.calc_force_density(0.0, 1e-7)
    orc.force(x)
    err = scaled_error(dh.get("b_int"), orc)
    say(f"  force scaled error {err:.3e}")
    assert err <= FP64_TOL
    dh.close()


def test_sets_line_factors(need_gpu, oracle):
    """Ten points on a line with E and 2E halves (test_bond_system_corrections.jl:126-146):
    averaged Lame parameters and bc at the interface."""
    say("esc sets line")
    X = np.zeros((3, 10))
    X[0] = np.arange(10.0)
    vol = np.array([1.1, 1.2, 1.3, 1.4, 1.5, 1.6, 1.7, 1.8, 1.9, 2.0])
    b = pd.Body(mat_of("BB"), X, vol)
    pd.point_set(b, "left", X[0] < 4.5)
    pd.point_set(b, "right", X[0] > 4.5)
    pd.material(b, "left", horizon=1.01, rho=8000.0, E=105e9, nu=0.25)
    pd.material(b, "right", horizon=1.01, rho=8000.0, E=210e9, nu=0.25)
    dh = pd.MI355XBodyDataHandler(b)
    assert dh.n_sets == 2
    orc = oracle_of(oracle, b, "BB")
    factors_match(dh, orc)
    dh.close()


@pytest.mark.parametrize("kind", ["BB", "OSB"])
def test_sets_with_correction_force(need_gpu, oracle, kind):
    say("esc + sets force", kind)
    pos, vol = pd.lattice_block(14, 12, 6, 1e-3, jitter_amplitude=0.05, seed=21)
    b = pd.Body(mat_of(kind), pos, vol)
    x0 = pos[0]
    pd.point_set(b, "l", x0 <= np.median(x0))
    pd.point_set(b, "r", x0 > np.median(x0))
    extra = {} if kind == "BB" else {"nu": 0.25}
    pd.material(b, "l", horizon=3.015e-3, rho=7850.0, E=210e9, epsilon_c=4e-4, **extra)
    pd.material(b, "r", horizon=3.015e-3, rho=2700.0, E=70e9, epsilon_c=1e-3, **extra)
    dh = pd.MI355XBodyDataHandler(b)
    orc = oracle_of(oracle, b, kind)
    factors_match(dh, orc)
    x = b.xyz + 6e-7 * np.random.default_rng(8).standard_normal(b.xyz.shape)
    dh.set("position", x)
    dh.calc_force_density(0.0, 1e-7)
    orc.force(x)
    err = scaled_error(dh.get("b_int"), orc)
    _, _, _, _, act = dh.bonds()
    say(f"  force scaled error {err:.3e}, broken {int(np.sum(orc.active == 0))}, "
        f"mismatches {int(np.sum(act != orc.active))}")
    assert err <= FP64_TOL
    assert np.array_equal(act, orc.active)
    dh.close()


def test_mixed_force_with_correction(need_gpu, oracle):
    say("esc mixed force")
    pos, vol = pd.lattice_block(16, 14, 6, 1e-3, jitter_amplitude=0.05, seed=5)
    b = pd.Body(mat_of("BB"), pos, vol)
    pd.material(b, horizon=3.015e-3, rho=2700.0, E=70e9)
    dh = pd.MI355XBodyDataHandler(b, precision="mixed")
    orc = oracle_of(oracle, b, "BB")
    x = b.xyz + 2e-7 * np.random.default_rng(6).standard_normal(b.xyz.shape)
    dh.set("position", x)
    dh.calc_force_density(0.0, 1e-7)
    orc.force(x)
    err = scaled_error(dh.get("b_int"), orc)
    say(f"  scaled error {err:.3e}")
    assert err <= 1e-5
    dh.close()


@pytest.mark.parametrize("kind", ["BB", "OSB"])
def test_vv_trajectory_with_correction(need_gpu, oracle, kind):
    say("esc vv", kind)
    pos, vol = pd.lattice_block(16, 16, 5, 1e-3, jitter_amplitude=0.05, seed=12)
    b = pd.Body(mat_of(kind), pos, vol)
    extra = {} if kind == "BB" else {"nu": 0.25}
    pd.material(b, horizon=3.015e-3, rho=7850.0, E=210e9, epsilon_c=5e-4, **extra)
    y = pos[1]
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
    say(f"  broken {int(np.sum(orc.active == 0))}, mismatches {int(np.sum(act != orc.active))}")
    assert np.array_equal(act, orc.active)
    dh.close()


def test_correction_rejected_for_cpd(need_gpu):
    """Correspondence materials take no surface correction: the C ABI refuses it."""
    from peridynamics import _lib
    pos, vol = pd.lattice_block(6, 6, 4, 1e-3)
    b = pd.Body(pd.CMaterial(), pos, vol)
    pd.material(b, horizon=3.015e-3, rho=7850.0, E=210e9, nu=0.25)
    p = pd.MI355XBodyDataHandler._pdx_params(b.point_params[0], b.mat)
    ctx = _lib.Context(0, _lib.PDX_FP64)
    ctx.set_points(b.xyz, b.volume, b.fail_permit.astype(np.uint8))
    ctx.set_material(b.mat.kind, p)
    ctx.set_surface_correction(1)
    with pytest.raises(pd.PdxError):
        ctx.build_bonds()
    ctx.close()


@pytest.mark.parametrize("kernel", ["const_one", "cubic_b_spline", "cubic_b_spline_norm"])
def test_osb_correction_with_influence_function(need_gpu, oracle, kernel):
    """OSBMaterial{EnergySurfaceCorrection} with an influence function other than
    linear_kernel: the strain energies of the factors weight the bonds by kernel(system,
    bond_id) (ordinary_state_based.jl:225), so mfactor, scfactor and the corrected force
    depend on it; all three against the oracle."""
    say("esc osb kernel", kernel)
    fn = {"const_one": pd.const_one_kernel, "cubic_b_spline": pd.cubic_b_spline_kernel,
          "cubic_b_spline_norm": pd.cubic_b_spline_kernel_norm}[kernel]
    pos, vol = pd.lattice_block(14, 12, 10, 1e-3, jitter_amplitude=0.05, seed=7)
    b = pd.Body(pd.OSBMaterial(kernel=fn, correction=pd.EnergySurfaceCorrection), pos, vol)
    pd.material(b, horizon=3.015e-3, rho=7850.0, E=210e9, nu=0.25)
    dh = pd.MI355XBodyDataHandler(b)
    orc = oracle.OracleSim(b.xyz, b.volume, "OSB", as_oracle_params(b.point_params[0], kernel=kernel),
                           fail_permit=b.fail_permit.astype(np.uint8), correction="energy")
    sc, _ = factors_match(dh, orc)
    # the correction differs from linear_kernel's
    lin = oracle.OracleSim(b.xyz, b.volume, "OSB", as_oracle_params(b.point_params[0]),
                           fail_permit=b.fail_permit.astype(np.uint8), correction="energy")
    assert np.max(np.abs(sc - lin.scf)) > 1e-6
    x = b.xyz + 1e-6 * np.random.default_rng(4).standard_normal(b.xyz.shape)
    dh.set("position", x)
    dh.calc_force_density(0.0, 1e-7)
    orc.force(x)
    err = scaled_error(dh.get("b_int"), orc)
    say(f"  force scaled error {err:.3e}")
    assert err <= FP64_TOL
    dh.close()
