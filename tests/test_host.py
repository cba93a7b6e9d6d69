"""Host-side logic of the reference interface (no GPU): material! parameter derivation,
solver constructors, point generators, BC bookkeeping, and the C-ABI library itself."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import peridynamics as pd
from peridynamics.materials import get_point_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_material_params_known_answers(golden):
    # test/unit/physics/test_bond_based.jl:25-152
    g = golden["material_params_bb"]
    exp = g["expect"]
    for kw in [g["kwargs"]] + g["equivalent_kwargs"]:
        p = get_point_params(pd.BBMaterial(), kw)
        for k, v in exp.items():
            assert getattr(p, k) == pytest.approx(v, rel=1.5e-8), (kw, k)
    for kw in g["rejected_kwargs"]:
        with pytest.raises(pd.ArgumentError):
            get_point_params(pd.BBMaterial(), kw)


def test_epsc_and_bond_constant_closed_forms():
    p = get_point_params(pd.BBMaterial(), dict(horizon=3.015e-3, rho=2700, E=70e9, Gc=100))
    K = 70e9 / (3 * (1 - 2 * 0.25))
    assert p.K == K
    assert p.epsc == math.sqrt(5.0 * 100 / (9.0 * K * 3.015e-3))
    assert p.bc == 18 * K / (math.pi * 3.015e-3 ** 4)
    q = get_point_params(pd.BBMaterial(), dict(horizon=1.0, rho=1, E=1, epsilon_c=0.01))
    assert q.Gc == pytest.approx(9 / 5 * q.K * 1.0 * 0.01 ** 2)
    with pytest.raises(pd.ArgumentError):
        get_point_params(pd.BBMaterial(), dict(horizon=1.0, rho=1, E=1, Gc=1, epsilon_c=0.01))


def test_osb_two_elastic_params_required():
    with pytest.raises(pd.ArgumentError):
        get_point_params(pd.OSBMaterial(), dict(horizon=1.0, rho=1, E=1))
    p = get_point_params(pd.OSBMaterial(), dict(horizon=1.0, rho=1, E=210e9, nu=0.3))
    assert p.G == pytest.approx(210e9 / 2.6)


def test_fail_permit_follows_fracture_params():
    pos, vol = pd.uniform_box(1, 1, 1, 0.5)
    b = pd.Body(pd.BBMaterial(), pos, vol)
    pd.material(b, horizon=2, rho=1, E=1)
    assert not b.fail_permit.any()  # no Gc -> failure prohibited (fracture.jl:147-155)
    pd.material(b, horizon=2, rho=1, E=1, Gc=1)
    assert b.fail_permit.all()
    pd.no_failure(b)
    assert not b.fail_permit.any()


def test_uniform_box_ordering():
    # point_generators.jl:45-60: x fastest, then y, then z, centred
    pos, vol = pd.uniform_box(1, 1, 1, 0.25)
    assert pos.shape == (3, 64)
    assert np.allclose(pos[:, 0], [-0.375, -0.375, -0.375])
    assert np.allclose(pos[:, 1], [-0.125, -0.375, -0.375])
    assert np.allclose(pos[:, 4], [-0.375, -0.125, -0.375])
    assert np.allclose(pos[:, 16], [-0.375, -0.375, -0.125])
    assert np.allclose(vol, 0.25 ** 3)
    assert abs(pos.mean()) < 1e-15


def test_jitter_is_seeded_and_bounded():
    pos, _ = pd.lattice_block(5, 4, 3, 1e-3)
    a = pd.jitter(pos, 1e-3, 0.05, seed=0x5EED)
    b = pd.jitter(pos, 1e-3, 0.05, seed=0x5EED)
    c = pd.jitter(pos, 1e-3, 0.05, seed=1)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.max(np.abs(a - pos)) <= 0.05e-3


def test_solver_constructors():
    with pytest.raises(pd.ArgumentError):
        pd.VelocityVerlet(time=1, steps=1)
    with pytest.raises(pd.ArgumentError):
        pd.VelocityVerlet()
    with pytest.raises(pd.ArgumentError):
        pd.VelocityVerlet(steps=1, safety_factor=1.0)
    nk = pd.NewtonKrylov(steps=10, stepsize=0.1)
    assert nk.n_steps == 10 and nk.end_time == pytest.approx(1.0)
    assert pd.NewtonKrylov(time=1.0, stepsize=0.3).n_steps == 4
    for bad in (dict(maxiter=0), dict(tol=0), dict(gmres_restart=0)):
        with pytest.raises(pd.ArgumentError):
            pd.NewtonKrylov(steps=1, **bad)


def test_boundary_condition_checks():
    pos, vol = pd.uniform_box(1, 1, 1, 0.5)
    b = pd.Body(pd.BBMaterial(), pos, vol)
    pd.material(b, horizon=2, rho=1, E=1)
    pd.point_set(b, "top", lambda p: p[2] > 0)
    assert len(b.point_sets["top"]) == 4
    pd.velocity_bc(lambda t: 1.0, b, "top", "z")
    with pytest.raises(pd.ArgumentError):  # conflict: same field, dim and points
        pd.velocity_bc(lambda t: 2.0, b, "all_points", 3)
    pd.velocity_bc(lambda p, t: p[0] * t, b, "top", 1)
    with pytest.raises(pd.ArgumentError):  # wrong argument names
        pd.velocity_bc(lambda x: 1.0, b, "top", 2)
    with pytest.raises(pd.ArgumentError):
        pd.displacement_bc(lambda t: 1.0, b, "top", 2)
    with pytest.raises(pd.ArgumentError):
        pd.velocity_bc(lambda t: 1.0, b, "top", 4)


def test_velocity_ic_checks_and_field():
    """velocity_ic! (initial_conditions.jl:178-193): a value or an f(p); conflicts with an
    existing condition of the same dimension on intersecting points, and functions with
    other arguments, are ArgumentErrors; apply_initial_conditions! (:86-94) sets the
    single-dim conditions first, then the position-dependent ones, skipping NaN."""
    pos, vol = pd.uniform_box(1, 1, 1, 0.5)
    b = pd.Body(pd.BBMaterial(), pos, vol)
    pd.material(b, horizon=2, rho=1, E=1)
    pd.point_set(b, "top", lambda p: p[2] > 0)
    pd.point_set(b, "left", lambda p: p[0] < 0)
    pd.velocity_ic(b, "top", "z", 3.0)
    with pytest.raises(pd.ArgumentError):  # same dimension on intersecting points
        pd.velocity_ic(b, "all_points", 3, 1.0)
    with pytest.raises(pd.ArgumentError):
        pd.velocity_ic(lambda p, t: 1.0, b, "left", "x")
    with pytest.raises(pd.ArgumentError):
        pd.velocity_ic(b, "nope", "x", 1.0)
    pd.velocity_ic(lambda p: np.nan if p[1] > 0 else 10.0 * p[0], b, "left", "x")
    pd.velocity_ic(b, "all_points", "y", -1.0)
    v = pd.body.initial_velocity(b)
    top, left = b.point_sets["top"], b.point_sets["left"]
    assert np.all(v[top, 2] == 3.0) and np.all(np.delete(v[:, 2], top) == 0)
    assert np.all(v[:, 1] == -1.0)
    X = b.position
    want = np.where(X[1, left] > 0, 0.0, 10.0 * X[0, left])
    np.testing.assert_array_equal(v[left, 0], want)
    assert np.all(np.delete(v[:, 0], left) == 0)


def test_eval_posdep_vectorised_and_scalar():
    from peridynamics.body import eval_posdep
    X = np.arange(12, dtype=float).reshape(3, 4)
    v = eval_posdep(lambda p, t: -10.0 * p[2] * t, X, 2.0)
    assert np.array_equal(v, -20.0 * X[2])
    w = eval_posdep(lambda p, t: 1.0 if p[0] > 1.5 else float("nan"), X, 0.0)
    assert np.isnan(w[0]) and np.isnan(w[1]) and w[2] == 1.0


# --------------------------------------------------------------------- the C ABI
def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "pdx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pdx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = pd.load_library()
    names = _declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    # the ctypes table covers the header one to one
    from peridynamics._lib import SIGNATURES
    assert sorted(SIGNATURES) == names


def test_library_version_and_null_context():
    lib = pd.load_library()
    assert b"gfx950" in lib.pdx_version()
    assert lib.pdx_last_error(None) == b"null context"
    assert lib.pdx_set_profiling(None, 1) == -1


def test_influence_functions_of_the_materials():
    """OSBMaterial / CMaterial(kernel=...) take the four influence functions of kernels.jl
    and pass their ids to the library (PDX_KERNEL_* of include/pdx.h); another function is
    an ArgumentError (the device evaluates these four only); each combines with the surface
    correction."""
    from peridynamics.materials import KERNEL_IDS
    assert [KERNEL_IDS[k] for k in (pd.linear_kernel, pd.const_one_kernel,
                                    pd.cubic_b_spline_kernel, pd.cubic_b_spline_kernel_norm)] == [0, 1, 2, 3]
    for k in KERNEL_IDS:
        assert pd.OSBMaterial(kernel=k).kernel is k
        assert pd.CMaterial(kernel=k).kernel is k
    with pytest.raises(pd.ArgumentError):
        pd.OSBMaterial(kernel=lambda d, L: 1.0)
    with pytest.raises(pd.ArgumentError):
        pd.CMaterial(kernel=lambda d, L: 1.0)
    m = pd.OSBMaterial(kernel=pd.cubic_b_spline_kernel, correction=pd.EnergySurfaceCorrection)
    assert m.kernel is pd.cubic_b_spline_kernel and m.correction is pd.EnergySurfaceCorrection
    # the mirror's scalar functions agree with kernels.jl's test values
    assert pd.cubic_b_spline_kernel(1.0, 0.5) == pytest.approx(1 / 6)
    assert pd.cubic_b_spline_kernel_norm(1.0, 0.75) == pytest.approx(8 / np.pi * 2 * 0.25 ** 3)
    assert pd.const_one_kernel(2.0, 1.0) == 1 and pd.linear_kernel(2.0, 1.0) == 2.0
