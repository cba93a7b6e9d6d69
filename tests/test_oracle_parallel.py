"""The oracle's OSB and CPD loops run in two phases (per point in parallel, then the
per-point gather of the bond forces over the reverse bond index).  They must give the serial
scatter loop's numbers bit for bit, for any thread count:

- thread-count invariance at a strained state with broken bonds, killed points, several
  parameter sets and the surface correction;
- the gather order: a plain-Python serial scatter (ordinary_state_based.jl:128-149, the
  reference's own loop order: b_int[i] += t V_j and b_int[j] -= t V_i while i runs
  upwards) from the oracle's per-point weighted volume and dilatation reproduces its b_int
  bit for bit.  Summing a point's contributions in another order (own bonds first, or the
  reverse owners) changes the last bits and fails it."""
import numpy as np
import pytest

E, NU = 70e9, 0.25
G, K = E / (2 * (1 + NU)), E / (3 * (1 - 2 * NU))
LAM = E * NU / ((1 + NU) * (1 - 2 * NU))


def lattice(nx, ny, nz, seed=1):
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"),
                 -1).reshape(-1, 3).astype(np.float64)
    return g + rng.uniform(-0.05, 0.05, g.shape), np.ones(len(g))


def params(delta=3.015, **kw):
    d = dict(delta=delta, rho=2700.0, E=E, nu=NU, G=G, K=K, mu=G, bc=18 * K / (np.pi * delta ** 4),
             epsc=0.004, Gc=1.0, Cs=0.5, maxdmg=0.3, model="SVK")
    d["lambda"] = LAM
    d.update(kw)
    return d


def strained(X, seed=2, amp=0.008):
    rng = np.random.default_rng(seed)
    x = X + rng.standard_normal(X.shape) * amp
    x[X[:, 0] > X[:, 0].max() - 2.5] += np.array([0.1, 0.0, 0.0])
    return x


CASES = [("OSB", {}, False, None), ("OSB", {}, True, None), ("OSB", {}, False, "energy"),
         ("OSB", {}, True, "energy"), ("CPD", {"model": "SVK"}, False, None),
         ("CPD", {"model": "NHP", "zem": "Wan", "Cs": 0.0}, False, None),
         ("CPD", {"model": "SVK"}, True, None)]


@pytest.mark.parametrize("kind,extra,sets,corr", CASES)
def test_thread_count_invariance(oracle, kind, extra, sets, corr):
    X, vol = lattice(12, 10, 8)
    x = strained(X)
    out = []
    for nt in (1, 3, 8):
        if sets:
            p2 = params(**extra)
            p2.update(G=1.5 * G, K=1.3 * K, mu=1.5 * G, bc=1.4 * p2["bc"], epsc=0.006)
            p2["lambda"] = 1.2 * LAM
            ps = (X[:, 0] > 6).astype(int)
            sim = oracle.OracleSim(X, vol, kind, [params(**extra), p2], point_set=ps,
                                   nthreads=nt, correction=corr)
        else:
            sim = oracle.OracleSim(X, vol, kind, params(**extra), nthreads=nt, correction=corr)
        sim.force(x)
        out.append((sim.b_int.copy(), sim.b_abs.copy(), sim.damage.copy(), sim.active.copy(),
                    sim.n_active.copy(), sim.cauchy.copy()))
    broken = int(np.count_nonzero(out[0][3] == 0))
    assert broken > 0
    if kind == "CPD":
        assert np.count_nonzero(out[0][4] == 0) > 0  # killed points
    for o in out[1:]:
        for a, b in zip(out[0], o):
            np.testing.assert_array_equal(a, b)


def test_gather_reproduces_serial_scatter(oracle):
    X, vol = lattice(5, 4, 4, seed=4)
    vol = vol * np.linspace(0.9, 1.1, len(vol))
    p = params(delta=1.8, epsc=0.02)
    sim = oracle.OracleSim(X, vol, "OSB", p, nthreads=4)
    x = strained(X, seed=5, amp=0.01)
    sim.force(x)
    rp, nbr, L, om, act = sim.row_ptr, sim.nbr, sim.L, sim.omega, sim.active
    b = [[0.0, 0.0, 0.0] for _ in range(sim.n)]
    for i in range(sim.n):
        wvol = float(sim.wvol[i])
        if wvol == 0.0:
            continue
        dil = float(sim.dil[i])
        c1 = (15.0 * p["G"]) / wvol
        c2 = dil * ((3.0 * p["K"]) / wvol - c1 / 3.0)
        for q in range(rp[i], rp[i + 1]):
            j = int(nbr[q])
            d = [float(x[j, k]) - float(x[i, k]) for k in range(3)]
            l = float(sim.bond_length[q])
            w = float(om[q]) if act[q] else 0.0
            s = (w * (c2 * float(L[q]) + c1 * (l - float(L[q])))) / l
            for k in range(3):
                t = s * d[k]
                b[i][k] += t * float(vol[j])
                b[j][k] += -t * float(vol[i])
    np.testing.assert_array_equal(sim.b_int, np.array(b))
