"""The oracle's EnergySurfaceCorrection (bond_system_corrections.jl), pinned on the
reference's own fixtures (the two-point scfactor 3.1808625617603665 and force of
test_bond_based.jl:181-211, the mfactor/scfactor ranges of a 1 m box of
test_bond_system_corrections.jl), on closed forms and on symmetries.

What a plausible mistake does here:
- a wrong stretch energy (e.g. 1/8 lambda (l^2-1)^2, the finite-strain form) or a dropped
  0.25 in the bond energy moves the two-point factor off 3.18086...;
- a transposed direction (mfactor[d] for the wrong axis) breaks the axis-bond identities
  scfactor = mean mfactor of the bond's axis;
- a per-point rather than per-bond factor, or one that is not symmetric in i and j,
  breaks momentum conservation of the corrected force."""
import math

import numpy as np
import pytest

from conftest import as_oracle_params


def params(kind, E=210e9, delta=1.5, rho=1.0, nu=0.25):
    import peridynamics as pd
    mat = pd.BBMaterial() if kind == "BB" else pd.OSBMaterial()
    kw = dict(horizon=delta, rho=rho, E=E)
    if kind != "BB":
        kw["nu"] = nu
    return as_oracle_params(pd.materials.get_point_params(mat, kw))


def test_two_point_fixture(oracle, golden):
    g = golden["bb_surface_correction_two_points"]
    X = np.array(g["position"]).T.copy()
    orc = oracle.OracleSim(X, np.array(g["volume"]), "BB", params("BB", g["E"], g["horizon"]),
                           fail_permit=np.zeros(2, np.uint8), correction="energy")
    sc = g["scfactor"]
    # the reference checks it with isapprox's default rtol = sqrt(eps); the closed form
    # below pins it exactly: (lambda / 2 + mu) / (bc / 4) = 0.6 E pi delta^4 / 3 E = 1.0125 pi
    assert np.allclose(orc.scf, sc, rtol=math.sqrt(np.finfo(float).eps), atol=0)
    assert np.allclose(orc.scf, 1.0125 * math.pi, rtol=4e-16, atol=0)
    # closed form: psi_theory / psi_pd along x for two points at distance 1 (nu = 1/4, so
    # lambda = mu = 0.4 E, bc = 18 K / (pi delta^4) with K = E / 1.5)
    E, d = g["E"], g["horizon"]
    st = 1.001 - 1.0
    psi_t = (0.5 * 0.4 * E + 0.4 * E) * st * st
    eps = (1.001 - 1.0) / 1.0
    psi_pd = 0.25 * (18 * E / 1.5 / (math.pi * d ** 4)) * eps * eps * 1.0 * 1.0
    assert orc.mfactor[0, 0] == pytest.approx(psi_t / psi_pd, rel=1e-12)
    assert np.all(np.isinf(orc.mfactor[:, 1:]))  # no bond stretches under y or z stretch
    x = orc.X.copy()
    x[1, 0] = g["moved_x_of_point_1"]
    b = orc.force(x)
    b12 = sc * 18 * E / (3 * (1 - 2 * 0.25)) / (math.pi * d ** 4) * 1.0015 * 0.0015 / 1.0015 * 1.0
    np.testing.assert_allclose(b[:, 0], [b12, -b12], rtol=1e-9)
    assert np.all(b[:, 1:] == 0)


@pytest.mark.parametrize("kind", ["BB", "OSB"])
def test_box_ranges(oracle, golden, kind):
    import peridynamics as pd
    g = golden["surface_correction_box"]
    pos, vol = pd.uniform_box(*g["uniform_box"])
    dx = g["uniform_box"][3]
    p = params(kind, g["E"], g["horizon_over_dx"] * dx, g["rho"], g["nu"])
    orc = oracle.OracleSim(pos.T.copy(), vol, kind, p, fail_permit=np.zeros(len(vol), np.uint8),
                           correction="energy")
    r = g[kind]
    for d in range(3):
        assert abs(orc.mfactor[:, d].min() - r["mfactor_min"][0]) <= r["mfactor_min"][1]
        assert abs(orc.mfactor[:, d].max() - r["mfactor_max"][0]) <= r["mfactor_max"][1]
    assert abs(orc.scf.min() - r["scfactor_min"][0]) <= r["scfactor_min"][1]
    assert abs(orc.scf.max() - r["scfactor_max"][0]) <= r["scfactor_max"][1]
    # the box is symmetric about its centre: the factors of mirrored points agree
    X = orc.X
    for d in range(3):
        m = X.copy()
        m[:, d] *= -1
        order = np.lexsort(np.round(X, 9).T)
        morder = np.lexsort(np.round(m, 9).T)
        np.testing.assert_allclose(orc.mfactor[order], orc.mfactor[morder], rtol=1e-12)
    # bonds along an axis take the mean mfactor of that axis
    owner = np.repeat(np.arange(orc.n), np.diff(orc.row_ptr))
    D = X[orc.nbr] - X[owner]
    for d in range(3):
        others = [e for e in range(3) if e != d]
        on_axis = np.all(np.abs(D[:, others]) <= 1e-10, axis=1)
        assert on_axis.any()
        mean = (orc.mfactor[owner, d] + orc.mfactor[orc.nbr, d]) / 2
        np.testing.assert_allclose(orc.scf[on_axis], mean[on_axis], rtol=1e-14)


@pytest.mark.parametrize("kind", ["BB", "OSB"])
def test_corrected_force_conserves_momentum(oracle, kind):
    """scfactor is symmetric in the two ends of a bond, so the corrected force densities still
    sum to zero over the body (weighted by the volumes)."""
    import peridynamics as pd
    pos, vol = pd.lattice_block(9, 8, 6, 1e-3, jitter_amplitude=0.05, seed=9)
    vol = vol * np.random.default_rng(4).uniform(0.8, 1.2, len(vol))
    p = params(kind, 210e9, 3.015e-3, 7850.0)
    orc = oracle.OracleSim(pos.T.copy(), vol, kind, p, fail_permit=np.zeros(len(vol), np.uint8),
                           correction="energy")
    owner = np.repeat(np.arange(orc.n), np.diff(orc.row_ptr))
    key = {(int(i), int(j)): k for k, (i, j) in enumerate(zip(owner, orc.nbr))}
    rev = np.array([key[(int(j), int(i))] for i, j in zip(owner, orc.nbr)])
    assert np.array_equal(orc.scf, orc.scf[rev])
    assert orc.scf.max() > 1.2  # the surface is corrected
    x = orc.X + 5e-7 * np.random.default_rng(5).standard_normal(orc.X.shape)
    b = orc.force(x)
    total = (b * orc.vol[:, None]).sum(0)
    scale = (np.abs(b) * orc.vol[:, None]).sum(0)
    assert np.all(np.abs(total) <= 1e-12 * scale)


def test_sets_lame_average_line(oracle):
    """Ten points on a line, horizon 1.01 (nearest neighbours only), E and 2E halves: the
    interface points 4 and 5 (0-based) average the Lame parameters over their bonds
    (test_bond_system_corrections.jl:126-184), and their mfactor along x is the closed form
    psi_t(averaged lambda, mu) / psi_pd(averaged bc)."""
    X = np.zeros((10, 3))
    X[:, 0] = np.arange(10.0)
    vol = np.array([1.1, 1.2, 1.3, 1.4, 1.5, 1.6, 1.7, 1.8, 1.9, 2.0])
    pa = params("BB", 105e9, 1.01, 8000.0)
    pb = params("BB", 210e9, 1.01, 8000.0)
    sets = [pa, pb]
    ps = np.array([0] * 5 + [1] * 5)
    orc = oracle.OracleSim(X, vol, "BB", sets, fail_permit=np.zeros(10, np.uint8),
                           point_set=ps, correction="energy")
    st = 1.001 - 1.0
    P = [sets[k] for k in ps]

    def psi_pd(i):
        tot = 0.0
        for j in (i - 1, i + 1):
            if 0 <= j < 10:
                bc = (P[i]["bc"] + P[j]["bc"]) / 2
                tot += 0.25 * bc * st * st * 1.0 * vol[j]
        return tot

    def lame(i):
        nb = [j for j in (i - 1, i + 1) if 0 <= j < 10]
        la = sum((P[i]["lambda"] + P[j]["lambda"]) / 2 for j in nb) / len(nb)
        mu = sum((P[i]["mu"] + P[j]["mu"]) / 2 for j in nb) / len(nb)
        return la, mu

    for i in (0, 3, 4, 5, 6, 9):
        la, mu = lame(i)
        expect = (0.5 * la + mu) * st * st / psi_pd(i)
        assert orc.mfactor[i, 0] == pytest.approx(expect, rel=1e-12)
    la4, _ = lame(4)
    assert la4 == pytest.approx((P[3]["lambda"] + 2 * P[4]["lambda"] + P[5]["lambda"]) / 4, rel=1e-15)
