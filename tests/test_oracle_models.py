"""The oracle's constitutive models (first_piola_kirchhoff of constitutive_models.jl:33-255:
SaintVenantKirchhoff, LinearElastic, NeoHooke, NeoHookePenalty), pinned as the reference's
test/unit/physics/test_constitutive_models.jl pins them: zero stress at F = I, P = dPsi/dF
(here by central differences of the textbook strain energies, the reference uses
ForwardDiff), a symmetric Cauchy stress, positive stress under tension, LinearElastic equal
to Saint Venant-Kirchhoff, and the closed form of pure shear.

A transposed product (S F instead of F S), a dropped term of S or the wrong Voigt slot for a
shear component fails the energy-derivative check."""
import numpy as np
import pytest

E_MOD, NU = 210e9, 0.3
LAM = E_MOD * NU / ((1 + NU) * (1 - 2 * NU))
MU = E_MOD / (2 * (1 + NU))
K_MOD = E_MOD / (3 * (1 - 2 * NU))
F_TEST = np.array([[1.02, 0.013, -0.004], [0.006, 0.985, 0.011], [-0.009, 0.007, 1.017]])


def psi(model, F):
    """strain_energy_density of each model (constitutive_models.jl)."""
    C = F.T @ F
    J = np.linalg.det(F)
    if model in ("SVK", "LE"):
        E = 0.5 * (C - np.eye(3))
        return 0.5 * LAM * np.trace(E) ** 2 + MU * np.trace(E @ E)
    if model == "NH":
        return 0.5 * MU * (np.trace(C) - 3) - MU * np.log(J) + 0.5 * LAM * np.log(J) ** 2
    return 0.5 * MU * (np.trace(C) * J ** (-2 / 3) - 3) + K_MOD / 8 * (J ** 2 + J ** -2 - 2)


def P_of(oracle, model, F):
    return oracle.pk1(model, F, LAM, MU, MU, K_MOD, NU)


MODELS = ["SVK", "LE", "NH", "NHP"]


@pytest.mark.parametrize("model", MODELS)
def test_zero_stress_at_identity(oracle, model):
    assert np.max(np.abs(P_of(oracle, model, np.eye(3)))) <= 1e-6  # Pa, of ~1e11 moduli


@pytest.mark.parametrize("model", MODELS)
def test_pk1_is_the_energy_derivative(oracle, model):
    P = P_of(oracle, model, F_TEST)
    h = 1e-6
    dpsi = np.zeros((3, 3))
    for r in range(3):
        for c in range(3):
            Fp, Fm = F_TEST.copy(), F_TEST.copy()
            Fp[r, c] += h
            Fm[r, c] -= h
            dpsi[r, c] = (psi(model, Fp) - psi(model, Fm)) / (2 * h)
    np.testing.assert_allclose(P, dpsi, rtol=0, atol=1e-6 * np.abs(P).max())


@pytest.mark.parametrize("model", MODELS)
def test_cauchy_symmetric_and_tension_positive(oracle, model):
    P = P_of(oracle, model, F_TEST)
    sig = P @ F_TEST.T / np.linalg.det(F_TEST)
    np.testing.assert_allclose(sig, sig.T, rtol=0, atol=1e-12 * np.abs(sig).max())
    Pt = P_of(oracle, model, np.diag([1.001, 1.0, 1.0]))
    assert Pt[0, 0] > 0


def test_linear_elastic_equals_svk(oracle):
    np.testing.assert_allclose(P_of(oracle, "LE", F_TEST), P_of(oracle, "SVK", F_TEST),
                               rtol=1e-13, atol=0)


@pytest.mark.parametrize("model", ["SVK", "LE"])
def test_pure_shear_closed_form(oracle, model):
    """F = [1 g 0; 0 1 0; 0 0 1]: E = [0 g/2 0; g/2 g^2/2 0; 0 0 0], S = lam tr(E) I + 2 mu E.
    E22 = (1 + g^2 - 1) / 2 carries the rounding of 1 + g^2 (relative ~1e-10 here), so the
    comparison is the reference's isapprox (rtol sqrt(eps))."""
    g = 1e-3
    F = np.array([[1.0, g, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    E = np.array([[0.0, g / 2, 0.0], [g / 2, g * g / 2, 0.0], [0.0, 0.0, 0.0]])
    S = LAM * np.trace(E) * np.eye(3) + 2 * MU * E
    np.testing.assert_allclose(P_of(oracle, model, F), F @ S, rtol=1.5e-8,
                               atol=1.5e-8 * np.abs(F @ S).max())
