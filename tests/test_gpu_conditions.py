"""Initial and time-dependent conditions on the HIP path, against the oracle.

- velocity_ic! (initial_conditions.jl:178-193, applied at chunk creation by
  apply_initial_conditions!, threads_body_data_handler.jl:55): a constant velocity on one
  set and a position-dependent one (NaN leaves a point alone, apply_ic! :71-83) on another,
  then 40 Velocity Verlet steps.
- forcedensity_bc! with f(t) and f(p, t) (boundary_conditions.jl:559-575) on b_ext during
  Velocity Verlet, applied by apply_boundary_conditions! after each half step.
Bars: the trajectory bars of the parity tests (displacement 1e-9, velocity 1e-8 normwise)."""
import sys

import numpy as np
import pytest

import peridynamics as pd
from conftest import as_oracle_params

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def normwise(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def block(kind="BB"):
    pos, vol = pd.lattice_block(16, 12, 6, 1e-3, jitter_amplitude=0.05, seed=17)
    b = pd.Body({"BB": pd.BBMaterial, "OSB": pd.OSBMaterial}[kind](), pos, vol)
    kw = dict(horizon=3.015e-3, rho=7850.0, E=210e9, epsilon_c=1e-3)
    if kind != "BB":
        kw["nu"] = 0.25
    pd.material(b, **kw)
    x = pos[0]
    pd.point_set(b, "left", x < x.min() + 3.5e-3)
    pd.point_set(b, "right", x > x.max() - 3.5e-3)
    return b


def oracle_of(oracle, b, kind):
    return oracle.OracleSim(b.xyz, b.volume, kind, as_oracle_params(b.point_params[0]),
                            fail_permit=b.fail_permit.astype(np.uint8))


def run_both(oracle, b, kind, steps, bcs=(), fd_bcs=()):
    vv = pd.VelocityVerlet(steps=steps)
    dh = pd.MI355XBodyDataHandler(b)
    dh.init_time_solver(vv)
    dh.initialize(vv)
    dh.solve(vv)
    orc = oracle_of(oracle, b, kind)
    orc.v[...] = pd.body.initial_velocity(b)
    orc.force()
    orc.vv_run(steps, vv.dt, bcs, fd_bcs=fd_bcs)
    return dh, orc


@pytest.mark.parametrize("kind", ["BB", "OSB"])
def test_velocity_ic_trajectory(need_gpu, oracle, kind):
    say("velocity_ic", kind)
    b = block(kind)
    pd.velocity_ic(b, "left", "x", -0.8)
    # NaN for the lower half of the set: those points keep velocity 0
    pd.velocity_ic(lambda p: np.nan if p[1] < 0 else 40.0 * p[2], b, "right", "y")
    v0 = pd.body.initial_velocity(b)
    right = b.point_sets["right"]
    assert np.all(v0[b.point_sets["left"], 0] == -0.8)
    assert np.all(v0[right[b.position[1, right] < 0], 1] == 0)
    assert np.any(v0[right, 1] != 0)
    dh, orc = run_both(oracle, b, kind, 40)
    for f, ref, tol in (("displacement", orc.u, 1e-9), ("velocity", orc.v, 1e-8)):
        err = normwise(dh.get(f), ref)
        say(f"  {f} normwise error {err:.3e}")
        assert err < tol
    assert np.abs(orc.u).max() > 1e-7  # the initial velocity moved the body


def test_velocity_ic_is_set_before_step_one(need_gpu):
    """The data handler holds the initial velocity before any step (and the velocity of
    the points without a condition stays zero)."""
    say("velocity_ic field")
    b = block()
    pd.velocity_ic(b, "left", 3, 2.5)
    dh = pd.MI355XBodyDataHandler(b)
    np.testing.assert_array_equal(dh.get("velocity"), pd.body.initial_velocity(b))


@pytest.mark.parametrize("kind", ["BB", "OSB"])
def test_forcedensity_bc_trajectory(need_gpu, oracle, kind):
    say("forcedensity_bc vv", kind)
    b = block(kind)
    pd.forcedensity_bc(lambda t: 7e7 * min(t / 2e-6, 1.0), b, "right", "x")
    pd.forcedensity_bc(lambda p, t: -5e9 * p[2] * t / 1e-6, b, "left", "z")
    bcs = []
    fd = [("sd", b.point_sets["right"], 0, lambda t: 7e7 * min(t / 2e-6, 1.0)),
          ("pd", b.point_sets["left"], 2, lambda p, t: -5e9 * p[2] * t / 1e-6)]
    dh, orc = run_both(oracle, b, kind, 40, bcs, fd)
    for f, ref, tol in (("displacement", orc.u, 1e-9), ("velocity", orc.v, 1e-8)):
        err = normwise(dh.get(f), ref)
        say(f"  {f} normwise error {err:.3e}")
        assert err < tol
    assert np.abs(orc.u).max() > 1e-9
