"""The export pipeline on the device (pdx_export_begin / pdx_export_wait) and the VTK files
a Job writes through it (export_results / export_reference_results,
src/core/threads_body_data_handler.jl:197-212).  The snapshot is copied while the next
steps run, so every exported array is compared bit for bit with a run that reads the same
fields synchronously at the same steps."""
import os
import sys

import numpy as np
import pytest

import peridynamics as pd

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def tension_body(nx=20, ny=20, nz=5, kind="BB"):
    pos, vol = pd.lattice_block(nx, ny, nz, 1e-3, jitter_amplitude=0.05, seed=3)
    b = pd.Body(pd.BBMaterial() if kind == "BB" else pd.OSBMaterial(), pos, vol)
    kw = dict(horizon=3.015e-3, rho=7850.0, E=210e9, epsilon_c=5e-4)
    if kind != "BB":
        kw["nu"] = 0.25
    pd.material(b, **kw)
    y = pos[1]
    pd.point_set(b, "top", y > y.max() - 2.5e-3)
    pd.point_set(b, "bottom", y < y.min() + 2.5e-3)
    pd.velocity_bc(lambda t: 2.0, b, "top", "y")
    pd.velocity_bc(lambda t: -2.0, b, "bottom", "y")
    return b


def test_begin_wait_semantics(need_gpu):
    say("export begin/wait")
    b = tension_body()
    vv = pd.VelocityVerlet(steps=20)
    dh = pd.MI355XBodyDataHandler(b)
    dh.init_time_solver(vv)
    dh.initialize(vv)
    names = ["position", "displacement", "damage", "n_active_bonds"]
    with pytest.raises(pd.PdxError):
        dh.ctx.export_wait()  # nothing pending
    dh.ctx.vv_run(1, 10, vv.dt, dh.bc_values(range(1, 11), vv.dt))
    ref0 = {f: dh.get(f).copy() for f in names}
    dh.ctx.export_begin(names, 10)
    dh.ctx.vv_run(11, 5, vv.dt, dh.bc_values(range(11, 16), vv.dt))
    ref1 = {f: dh.get(f).copy() for f in names}
    dh.ctx.export_begin(names, 15)
    with pytest.raises(pd.PdxError):
        dh.ctx.export_begin(names, 99)  # two pending
    tag, slot, arr = dh.ctx.export_wait()
    assert tag == 10 and slot == 0
    for f in names:
        assert np.array_equal(arr[f], ref0[f]), f
    tag, slot, arr = dh.ctx.export_wait()
    assert tag == 15 and slot == 1
    for f in names:
        assert np.array_equal(arr[f], ref1[f]), f
    dh.close()


@pytest.mark.parametrize("kind", ["BB", "OSB"])
def test_job_vtk_files_match_fields(need_gpu, tmp_path, kind):
    say("job vtk", kind)
    fields = ("displacement", "damage", "velocity", "b_int")
    # the same job twice: fields read synchronously into memory, then VTK files
    b = tension_body(kind=kind)
    job = pd.Job(b, pd.VelocityVerlet(steps=60), freq=20, fields=fields)
    dh = pd.submit(job)
    ref = {r["step"]: r for r in job.results}
    dh.close()
    b = tension_body(kind=kind)
    root = str(tmp_path / "res")
    job2 = pd.Job(b, pd.VelocityVerlet(steps=60), path=root, freq=20, fields=fields)
    dh2 = pd.submit(job2)
    dt = job2.solver.dt
    files = sorted(os.listdir(os.path.join(root, "vtk")))
    assert [f for f in files if f.endswith(".pvtu")] == [
        "timestep_0.pvtu", "timestep_20.pvtu", "timestep_40.pvtu", "timestep_60.pvtu"]
    X = b.xyz
    for n in (20, 40, 60):
        r = pd.read_vtk(os.path.join(root, "vtk", f"timestep_{n}.pvtu"))
        assert r["time"][0] == n * dt
        for f in fields:
            assert np.array_equal(r[f], ref[n][f]), (n, f)
        # the points are the current positions x = X + u (updated separately, so to rounding)
        assert np.allclose(r["position"] - X, r["displacement"], rtol=0, atol=1e-15)
    r0 = pd.read_vtk(os.path.join(root, "vtk", "timestep_0.pvtu"))
    assert np.array_equal(r0["position"], X) and np.all(r0["displacement"] == 0)
    assert np.any(r["damage"] > 0)
    dh2.close()


@pytest.mark.parametrize("kind", ["BB", "OSB"])
def test_exported_fields_match_oracle(need_gpu, oracle, tmp_path, kind):
    """The VTK files of a Job against the oracle's state at the export steps
    (export_results every freq steps, threads_body_data_handler.jl:197-212; the reference
    ships no .vtu fixtures, so the oracle's trajectory is the independent data): displacement
    and velocity within the trajectory bars (1e-9, 1e-8 normwise), damage bit for bit, b_int
    within 1e-8 of each point's bond-sum rounding scale, the time of each file exact."""
    from conftest import as_oracle_params
    say("vtk vs oracle", kind)
    fields = ("displacement", "velocity", "damage", "b_int")
    b = tension_body(kind=kind)
    root = str(tmp_path / "res")
    job = pd.Job(b, pd.VelocityVerlet(steps=60), path=root, freq=20, fields=fields)
    dh = pd.submit(job)
    dt = job.solver.dt
    dh.close()
    orc = oracle.OracleSim(b.xyz, b.volume, kind, as_oracle_params(b.point_params[0]),
                           fail_permit=b.fail_permit.astype(np.uint8))
    assert dt == orc.stable_dt(0.7)
    bcs = [("sd", b.point_sets["top"], 1, lambda t: 2.0),
           ("sd", b.point_sets["bottom"], 1, lambda t: -2.0)]
    orc.force()
    for n in (20, 40, 60):
        orc.vv_run(20, dt, bcs, first_step=n - 19)
        r = pd.read_vtk(os.path.join(root, "vtk", f"timestep_{n}.pvtu"))
        assert r["time"][0] == n * dt
        eu = float(np.max(np.abs(r["displacement"] - orc.u)) / np.max(np.abs(orc.u)))
        ev = float(np.max(np.abs(r["velocity"] - orc.v)) / np.max(np.abs(orc.v)))
        scale = np.max(orc.b_abs, axis=1)
        m = scale > 0
        eb = float(np.max(np.max(np.abs(r["b_int"] - orc.b_int), axis=1)[m] / scale[m]))
        say(f"  step {n}: displacement {eu:.2e}, velocity {ev:.2e}, b_int {eb:.2e}, "
            f"damaged points {int(np.sum(orc.damage > 0))}")
        assert eu < 1e-9 and ev < 1e-8 and eb < 1e-8
        assert np.array_equal(r["damage"], orc.damage)
    assert np.any(orc.damage > 0)
