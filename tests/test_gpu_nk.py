"""Newton-Krylov on the device (libpdx) against the CPU oracle and the reference's closed
forms.  J*v is a central finite difference, so its parity bar is relative to ||J v|| at
1e-8 (the rounding of the two force sums divided by 2 eps); converged displacements agree
to the solver tolerance."""
import sys

import numpy as np
import pytest

import peridynamics as pd
from conftest import as_oracle_params

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def nk_block(nx=16, ny=16, nz=4, dx=1e-3, top_disp=1e-4):
    """Configuration-4 style block: bottom face fixed, top face displaced in y."""
    pos, vol = pd.lattice_block(nx, ny, nz, dx, jitter_amplitude=0.02, seed=11)
    b = pd.Body(pd.BBMaterial(), pos, vol)
    pd.material(b, horizon=3.015 * dx, rho=2700.0, E=70e9)
    y = pos[1]
    pd.point_set(b, "bottom", y < y.min() + 1.5 * dx)
    pd.point_set(b, "top", y > y.max() - 1.5 * dx)
    for d in "xyz":
        pd.displacement_bc(lambda p: 0.0, b, "bottom", d)
    pd.displacement_bc(lambda p: top_disp, b, "top", "y")
    return b


def oracle_sim(oracle, b):
    p = b.point_params[0]
    return oracle.OracleSim(b.xyz, b.volume, "BB", as_oracle_params(p),
                            fail_permit=b.fail_permit.astype(np.uint8))


def disp_list(b, top_disp=1e-4):
    return ([(b.point_sets["bottom"], d, lambda p: 0.0) for d in range(3)] +
            [(b.point_sets["top"], 1, lambda p: top_disp)])


def test_residual_and_jv_parity(need_gpu, oracle):
    say("nk residual and jv")
    b = nk_block()
    dh = pd.MI355XBodyDataHandler(b)
    nk = pd.NewtonKrylov(steps=2)
    dh.init_newton_krylov(nk)
    orc = oracle_sim(oracle, b)
    rng = np.random.default_rng(5)
    u = 1e-6 * rng.standard_normal((b.n_points, 3))
    dh.set("displacement", u)
    dh.set("position", b.xyz + u)
    orc.u[...] = u
    orc.x[...] = b.xyz + u
    r, rn = dh.ctx.residual()
    cons = orc.constrained_dofs(disp_list(b))
    orc.force()
    ro = orc.residual(cons).reshape(-1, 3)
    scale = np.max(orc.b_abs * b.volume[:, None])
    err = np.max(np.abs(r - ro)) / scale
    say(f"  residual error {err:.2e} (of the bond-term scale)")
    assert err < 1e-12
    assert rn == pytest.approx(np.sqrt(np.sum(ro ** 2)), rel=1e-9)
    v = rng.standard_normal((b.n_points, 3))
    jv = dh.ctx.jv(v)
    jo = orc.jv(v.reshape(-1), cons).reshape(-1, 3)
    rel = np.max(np.abs(jv - jo)) / np.max(np.abs(jo))
    say(f"  jv error {rel:.2e} (normwise)")
    assert rel < 1e-8
    # constrained rows are identity
    cm = np.zeros(3 * b.n_points, bool)
    cm[cons] = True
    np.testing.assert_array_equal(jv.reshape(-1)[cm], v.reshape(-1)[cm])


class _env:
    """Environment variables set inside the block (the J v path switches are read when the
    Newton-Krylov state is set up, PDX_JV_TWO_KERNELS, and at each launch, PDX_JV_ONE_PASS)."""

    def __init__(self, **kv):
        self.kv = kv

    def __enter__(self):
        import os
        self.old = {k: os.environ.get(k) for k in self.kv}
        os.environ.update(self.kv)

    def __exit__(self, *exc):
        import os
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def test_fused_jv_equals_two_force_evaluations(need_gpu):
    """J v in one kernel against the two force evaluations of the step kernel at u + e v and
    u - e v: the same bits, for both one-kernel forms (both perturbed tables staged at once,
    and one after the other), at a strained state (a uniform stretch of 1e-3 along y plus a
    random field), for a v that is zero on the constrained rows and for one that is not.
    Newton-Krylov refuses materials that can fail, so the bonds never break inside J v."""
    say("nk fused jv vs two force evaluations")
    dx = 1e-3
    pos, vol = pd.lattice_block(24, 24, 6, dx, jitter_amplitude=0.02, seed=3)
    b = pd.Body(pd.BBMaterial(), pos, vol)
    pd.material(b, horizon=3.015 * dx, rho=2700.0, E=70e9)
    y = pos[1]
    pd.point_set(b, "bottom", y < y.min() + 1.5 * dx)
    for d in "xyz":
        pd.displacement_bc(lambda p: 0.0, b, "bottom", d)
    rng = np.random.default_rng(8)
    u = 1e-7 * rng.standard_normal((b.n_points, 3))
    u[:, 1] += 1e-3 * (y - y.min())
    v1 = rng.standard_normal((b.n_points, 3))
    v2 = v1.copy()
    v2[b.point_sets["bottom"]] = 0.0
    out = {}
    for name, env in (("one pass", {}), ("sequential", {"PDX_JV_ONE_PASS": "0"}),
                      ("two kernels", {"PDX_JV_TWO_KERNELS": "1"})):
        with _env(**env):
            dh = pd.MI355XBodyDataHandler(b)
            dh.init_newton_krylov(pd.NewtonKrylov(steps=1))
            dh.set("displacement", u)
            dh.set("position", b.xyz + u)
            out[name] = [dh.ctx.jv(v) for v in (v1, v2)]
            dh.close()
    for name in ("one pass", "sequential"):
        for a, c in zip(out[name], out["two kernels"]):
            assert np.all(np.isfinite(a)) and np.max(np.abs(a)) > 0
            np.testing.assert_array_equal(a, c, err_msg=name)


@pytest.mark.parametrize("precond", ["none", "jacobi"])
def test_nk_load_steps_match_oracle(need_gpu, oracle, precond):
    say("nk steps", precond)
    b = nk_block()
    nk = pd.NewtonKrylov(steps=2, tol=1e-4, maxiter=30, gmres_restart=30, gmres_maxiter=300,
                         precond=precond)
    dh = pd.MI355XBodyDataHandler(b)
    stats = dh.solve_nk(nk)
    for st in stats:
        say(f"  newton {st.newton_iters} krylov {st.krylov_iters} r {st.residual_first:.3e}"
            f" -> {st.residual_final:.3e}")
        assert st.residual_final < 1e-4
    orc = oracle_sim(oracle, b)
    for n in (1, 2):
        orc.nk_step(n, 2, disp_list(b), maxiter=30, tol=1e-4, gmres_restart=30, gmres_maxiter=300)
    u = dh.get("displacement")
    err = np.max(np.abs(u - orc.u)) / np.max(np.abs(orc.u))
    say(f"  displacement vs oracle {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("precond", ["jacobi", "two_level"])
def test_preconditioners_on_a_ragged_size(need_gpu, precond):
    """37x37x7 points: the last 256-point block of the per-point kernels is partly empty and
    its last wave only partly filled.  Both preconditioners' solves stay finite, reach the
    relative tolerance and land on the same displacement (1e-5 normwise: a residual of 1e-8
    relative leaves the displacement determined to the stiffness's condition times that)."""
    say("nk ragged size", precond)
    b = nk_block(37, 37, 7)
    assert b.n_points % 256 and b.n_points % 64
    nk = pd.NewtonKrylov(steps=2, tol=1e-30, rel_tol=1e-8, maxiter=30, gmres_restart=30,
                         gmres_maxiter=300, precond=precond)
    dh = pd.MI355XBodyDataHandler(b)
    stats = dh.solve_nk(nk)
    for st in stats:
        say(f"  newton {st.newton_iters} krylov {st.krylov_iters} r {st.residual_first:.3e}"
            f" -> {st.residual_final:.3e}")
        assert np.isfinite(st.residual_final) and st.residual_final < 1e-8 * st.residual_first
    u = dh.get("displacement")
    dh.close()
    assert np.all(np.isfinite(u))
    ref = _RAGGED.setdefault("u", u)
    assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 1e-5


_RAGGED = {}


def test_jacobi_reduces_krylov_iterations(need_gpu):
    say("nk jacobi vs plain")
    its = {}
    for pre in ("none", "jacobi"):
        b = nk_block(24, 24, 4)
        nk = pd.NewtonKrylov(steps=1, tol=1e-6, maxiter=30, gmres_restart=30, gmres_maxiter=2000,
                             gmres_reltol=1e-6, precond=pre)
        dh = pd.MI355XBodyDataHandler(b)
        st = dh.solve_nk(nk)[0]
        its[pre] = st.krylov_iters
        say(f"  {pre}: krylov {st.krylov_iters}, newton {st.newton_iters}, {st.seconds:.3f} s")
    assert its["jacobi"] < its["none"]


def test_nk_uniform_tension_closed_form(need_gpu):
    """test/simulation/test_analytic_solutions.jl:7-94 on the device: CMaterial bar, 297
    points, five load steps, elongation within 2 % of F l / (E A)."""
    say("nk uniform tension")
    l, w, h, Ftot, E = 1.0, 0.1, 0.1, 2e6, 200e9
    dx = 1 / 30
    pos, vol = pd.uniform_box(l + 3 * dx, w, h, dx, center=(-1.5 * dx, 0, 0))
    b = pd.Body(pd.CMaterial(zem=pd.ZEMSilling(Cs=0.5)), pos, vol)
    pd.material(b, horizon=3.015 * dx, rho=7850.0, E=E, nu=0.25)
    pd.point_set(b, "left", lambda p: p[0] < -l / 2)
    pd.point_set(b, "right", lambda p: p[0] > l / 2 - dx)
    b_right = Ftot / vol[b.point_sets["right"]].sum()
    for d in "xyz":
        pd.displacement_bc(lambda p: 0.0, b, "left", d)
    pd.forcedensity_bc(lambda p: b_right, b, "right", "x")
    nk = pd.NewtonKrylov(steps=5, tol=1e-3, maxiter=50)
    job = pd.Job(b, nk, fields=())
    dh = pd.submit(job)
    u = dh.get("displacement")
    elong = u[b.point_sets["right"], 0].mean()
    exact = Ftot / (E * w * h) * l
    say(f"  elongation {elong:.6e} exact {exact:.6e}")
    assert abs(elong - exact) / exact < 0.02


def test_nk_throws_at_maxiter(need_gpu):
    """test/unit/time_solvers/test_newton_krylov.jl:393-409."""
    say("nk maxiter")
    pos = np.array([[0.0, 1.0, 0.0, 0.0], [0.0, 0.0, 1.0, 0.0], [0.0, 0.0, 0.0, 1.0]])
    b = pd.Body(pd.BBMaterial(), pos, np.ones(4))
    pd.material(b, horizon=2, rho=1, E=1)
    pd.point_set(b, "top", [3])
    pd.point_set(b, "bottom", [0])
    pd.displacement_bc(lambda p: 0.1, b, "top", "y")
    for d in "xyz":
        pd.displacement_bc(lambda p: 0.0, b, "bottom", d)
    nk = pd.NewtonKrylov(steps=10, stepsize=0.1, maxiter=2, tol=1e-6)
    with pytest.raises(pd.NotConvergedError):
        pd.submit(pd.Job(b, nk, fields=()))


def _tension_bar():
    """The CMaterial bar of test_analytic_solutions.jl:7-94 (see the closed-form test)."""
    l, w, h, Ftot, E = 1.0, 0.1, 0.1, 2e6, 200e9
    dx = 1 / 30
    pos, vol = pd.uniform_box(l + 3 * dx, w, h, dx, center=(-1.5 * dx, 0, 0))
    b = pd.Body(pd.CMaterial(zem=pd.ZEMSilling(Cs=0.5)), pos, vol)
    pd.material(b, horizon=3.015 * dx, rho=7850.0, E=E, nu=0.25)
    pd.point_set(b, "left", lambda p: p[0] < -l / 2)
    pd.point_set(b, "right", lambda p: p[0] > l / 2 - dx)
    b_right = Ftot / vol[b.point_sets["right"]].sum()
    for d in "xyz":
        pd.displacement_bc(lambda p: 0.0, b, "left", d)
    pd.forcedensity_bc(lambda p: b_right, b, "right", "x")
    return b


@pytest.mark.parametrize("kind", ["BB", "CPD"])
def test_nk_in_mixed_context_uses_fp64_forces(need_gpu, kind):
    """A mixed-precision context solves Newton-Krylov with the FP64 force kernels (the
    finite-difference Jacobian products need them): its load steps equal those of an FP64
    context bit for bit."""
    say("nk mixed context", kind)
    out = {}
    for prec in ("fp64", "mixed"):
        if kind == "BB":
            b = nk_block()
            nk = pd.NewtonKrylov(steps=2, tol=1e-4, maxiter=30, gmres_restart=30,
                                 gmres_maxiter=300, precond="jacobi")
        else:
            b = _tension_bar()
            nk = pd.NewtonKrylov(steps=2, tol=1e-3, maxiter=50)
        dh = pd.MI355XBodyDataHandler(b, precision=prec)
        stats = dh.solve_nk(nk)
        out[prec] = (dh.get("displacement"), [(st.newton_iters, st.krylov_iters) for st in stats])
        say(f"  {prec}: {out[prec][1]}")
        dh.close()
    assert out["fp64"][1] == out["mixed"][1]
    assert np.array_equal(out["fp64"][0], out["mixed"][0])
