"""Pins the oracle's DynamicRelaxation on the reference's known answers
(test/unit/time_solvers/test_dynamic_relaxation.jl)."""
import numpy as np
import pytest

import peridynamics as pd
from conftest import as_oracle_params

TETRA = np.array([[0.0, 1.0, 0.0, 0.0], [0.0, 0.0, 0.0, 1.0], [0.0, 0.0, 1.0, 0.0]])


def sim(oracle, body):
    return oracle.OracleSim(body.xyz, body.volume, "BB", as_oracle_params(body.point_params[0]))


@pytest.mark.parametrize("E,expected", [(210e3, 4.2e6), (200e3, 4.0e6)])
def test_density_matrix(oracle, E, expected):
    b = pd.Body(pd.BBMaterial(), TETRA, np.ones(4))
    pd.material(b, horizon=1.5, rho=8e-6, E=E)
    np.testing.assert_allclose(sim(oracle, b).dr_density(1.0, 1.0), expected, rtol=1e-14)


def test_two_steps(oracle):
    b = pd.Body(pd.BBMaterial(), TETRA, np.ones(4))
    pd.material(b, horizon=3.0, rho=7.8e-6, E=200e3)
    o = sim(oracle, b)
    o.force()
    o.dr_run(2, 1.0, 1.0, fd_bcs=[("pd", [0, 1], 0, lambda X, t: 1e5 * X[0])])
    assert o.x[1, 0] == pytest.approx(1.05, rel=1e-14)


def test_damping(oracle):
    pos = np.zeros((3, 10))
    pos[0] = np.arange(10.0)
    b = pd.Body(pd.BBMaterial(), pos, np.ones(10))
    pd.material(b, horizon=1.5, rho=1.0, E=1.0, nu=0.25)
    o = sim(oracle, b)
    dm = o.dr_density(1.0, 1.0)
    assert np.all(dm > 0)
    o.u[...] = 1.0
    o.vh[...] = 1.0
    vh_old, bint_old = np.ones((10, 3)), np.zeros((10, 3))
    o.b_int[...] = -1e6
    assert o.dr_damping(1.0, dm, vh_old, bint_old) == 1.9
    o.b_int[...] = -1e-6
    cn1 = sum(1e-6 / d for d in np.repeat(dm, 3))
    assert o.dr_damping(1.0, dm, vh_old, bint_old) == pytest.approx(2 * np.sqrt(cn1 / 30), rel=1e-14)
    o.u[...] = 0.0
    assert o.dr_damping(1.0, dm, vh_old, bint_old) == 0.0
