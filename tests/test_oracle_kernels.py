"""The oracle's influence functions omega(delta, L) (src/discretization/kernels.jl), pinned
by the reference's own test values (test/unit/discretization/test_kernels.jl) and by what
the mathematics fixes: the normalised cubic B-spline integrates to 1 over the horizon
sphere, it is 3C/2 times the plain one (C = 8 / (pi delta^3)), both are C1 at xi = 1/2 and
vanish with zero slope at xi = 1.  A wrong coefficient, a swapped branch or a dropped power
fails one of these."""
import numpy as np
import pytest
from scipy import integrate

KERNELS = ["linear", "const_one", "cubic_b_spline", "cubic_b_spline_norm"]


def test_reference_test_values(oracle):
    f = oracle.influence
    # const_one_kernel(1.0, 2.0) == 1 ... ; linear_kernel(2.0, 1.0) = 2, (1.0, 2.0) = 0.5
    assert f("const_one", 1.0, [2.0, 0.1, 0.001]).tolist() == [1.0, 1.0, 1.0]
    assert f("linear", 2.0, [1.0])[0] == 2.0 and f("linear", 1.0, [2.0])[0] == 0.5
    # cubic_b_spline_kernel(1, L) at L = 0, delta: 0; 0.5, 0.25, 0.75: the two branches
    c = f("cubic_b_spline", 1.0, [0.0, 1.0, 0.5, 0.25, 0.75])
    assert abs(c[0]) <= np.finfo(float).eps and abs(c[1]) <= np.finfo(float).eps
    np.testing.assert_allclose(c[2:], [1 / 6, 2 / 3 - 0.25 + 0.0625, 4 / 3 - 3 + 2.25 - 0.5625],
                               rtol=1e-14)
    C = 8 / np.pi
    n = f("cubic_b_spline_norm", 1.0, [0.0, 1.0, 0.5, 0.25, 0.75])
    assert abs(n[0]) <= np.finfo(float).eps and abs(n[1]) <= np.finfo(float).eps
    np.testing.assert_allclose(n[2:], [C * 0.25, C * (1 - 0.375 + 0.09375), C * 2 * 0.015625],
                               rtol=1e-14)


@pytest.mark.parametrize("delta", [1.0, 3.015e-3])
def test_normalised_spline_integrates_to_one(oracle, delta):
    g = lambda r: oracle.influence("cubic_b_spline_norm", delta, np.array([r]))[0] * 4 * np.pi * r * r
    val, _ = integrate.quad(g, 0.0, delta, points=[0.5 * delta], epsabs=0, epsrel=1e-13)
    assert val == pytest.approx(1.0, rel=1e-12)


def test_normalised_is_scaled_plain_spline(oracle):
    delta = 2.5
    L = np.linspace(1e-3, 1.2 * delta, 997)
    C = 8.0 / (np.pi * delta ** 3)
    np.testing.assert_allclose(oracle.influence("cubic_b_spline_norm", delta, L),
                               1.5 * C * oracle.influence("cubic_b_spline", delta, L),
                               rtol=1e-13, atol=1e-15 * C)  # cancellation near xi = 1


@pytest.mark.parametrize("kernel", ["cubic_b_spline", "cubic_b_spline_norm"])
def test_spline_continuity(oracle, kernel):
    f = lambda x: oracle.influence(kernel, 1.0, np.array([x]))[0]
    h = 1e-6
    for x0 in (0.5, 1.0):
        assert f(x0 - h) == pytest.approx(f(x0 + h) if x0 < 1 else 0.0, abs=1e-5)
        dl = (f(x0) - f(x0 - h)) / h
        dr = (f(x0 + h) - f(x0)) / h
        assert dl == pytest.approx(dr, abs=1e-4)  # C1 (slope 0 at the support's edge)
    # outside the support and at L = 0
    assert f(1.0 + 1e-9) == 0.0 and f(0.0) == 0.0


def test_oracle_sim_uses_the_kernel(oracle):
    X = np.stack(np.meshgrid(np.arange(6.0), np.arange(5.0), np.arange(4.0), indexing="ij"),
                 -1).reshape(-1, 3)
    base = dict(delta=3.015, rho=1.0, E=1.0, nu=0.25, G=0.4, K=2 / 3, bc=1.0, epsc=1.0, Gc=1.0,
                mu=0.4)
    base["lambda"] = 0.4
    for k in KERNELS:
        sim = oracle.OracleSim(X, np.ones(len(X)), "OSB", dict(base, kernel=k))
        np.testing.assert_array_equal(sim.omega, oracle.influence(k, 3.015, sim.L))
