"""pdx_vv_run_noalloc through the C ABI: the Velocity Verlet run whose boundary-condition
value table is staged in the library's fixed device array instead of context memory, so
that the call allocates nothing.  Its results must equal pdx_vv_run's bit for bit (the same
kernels read the same values), on two calls in a row, the second with a smaller table, and
a run whose table exceeds the fixed array is refused with an error."""
import sys

import numpy as np
import pytest

import peridynamics as pd
from peridynamics._lib import PdxError

pytestmark = pytest.mark.gpu

FIELDS = ("displacement", "velocity", "acceleration", "damage", "n_active_bonds")


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def torsion_body():
    """A jittered bar with position-dependent velocity conditions on one end (one value per
    point and step) and fixed conditions on the other, bonds breaking from step ~20."""
    pos, vol = pd.lattice_block(18, 10, 8, 1e-3, jitter_amplitude=0.05, seed=21)
    b = pd.Body(pd.BBMaterial(), pos, vol)
    pd.material(b, horizon=3.015e-3, rho=7850.0, E=210e9, epsilon_c=4e-4)
    x0 = pos[0]
    pd.point_set(b, "end", x0 > x0.max() - 2.5e-3)
    pd.point_set(b, "fixed", x0 < x0.min() + 2.5e-3)
    pd.velocity_bc(lambda p, t: -40.0 * p[2] / 4e-3, b, "end", "y")
    pd.velocity_bc(lambda p, t: 40.0 * p[1] / 4e-3 + 5.0, b, "end", "z")
    pd.velocity_bc(lambda t: 0.0, b, "fixed", "x")
    return b


def make():
    b = torsion_body()
    vv = pd.VelocityVerlet(steps=200)
    dh = pd.MI355XBodyDataHandler(b)
    dh.init_time_solver(vv)
    dh.initialize(vv)
    return dh, vv


def state(dh):
    _, _, _, _, act = dh.bonds()
    return {f: dh.get(f) for f in FIELDS} | {"active": act}


def test_noalloc_run_equals_vv_run(need_gpu):
    say("vv_run_noalloc vs vv_run")
    ref, vv = make()
    new, _ = make()
    assert ref.n_src > 0
    dt = vv.dt
    first = 1
    for n_steps in (60, 7):  # the second call stages a smaller table over the first one's
        vals = ref.bc_values(range(first, first + n_steps), dt)
        say(f"  steps {first}..{first + n_steps - 1}, table of {vals.size} values")
        ref.ctx.vv_run(first, n_steps, dt, vals)
        new.ctx.vv_run(first, n_steps, dt, vals, noalloc=True)
        a, b = state(ref), state(new)
        for k in a:
            assert np.array_equal(a[k], b[k]), (n_steps, k)
        first += n_steps
    assert np.any(a["active"] == 0)  # the run broke bonds
    assert np.any(a["displacement"] != 0)
    ref.close()
    new.close()


def test_noalloc_refuses_a_table_beyond_its_array(need_gpu):
    say("vv_run_noalloc size limit")
    dh, vv = make()
    before = dh.get("displacement")
    limit = 1 << 19
    n_steps = limit // dh.n_src + 1
    vals = np.zeros((n_steps, dh.n_src))
    with pytest.raises(PdxError, match="exceed the fixed table"):
        dh.ctx.vv_run(1, n_steps, vv.dt, vals, noalloc=True)
    assert np.array_equal(dh.get("displacement"), before)
    # a table that just fits runs
    n_fit = limit // dh.n_src
    assert n_fit * dh.n_src <= limit
    dh.ctx.vv_run(1, 3, vv.dt, dh.bc_values(range(1, 4), vv.dt), noalloc=True)
    dh.close()
