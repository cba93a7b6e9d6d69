"""CPU oracle driver for the force-density hot path (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker or the timed CPU
baseline.  The product path (``libpdx.so`` behind ``include/pdx.h``) never calls it.

The numerical kernels live in ``pdx_oracle.c`` (plain C restatement of the
reference, one function per reference routine, file:line cited there).  This module
adds the drivers the reference runs around them:

* :func:`find_bonds`             bond_system.jl:96-132 (+ precracks :134-165)
* :class:`OracleSim.force`       calc_force_density!, bond_system.jl:285-297
* :meth:`OracleSim.vv_run`       solve!/verlet_timestep!, velocity_verlet.jl:165-197
* :meth:`OracleSim.stable_dt`    calc_stable_timestep, velocity_verlet.jl:136-163
* :meth:`OracleSim.nk_step`      newton_krylov_step!, newton_krylov.jl:298-332
* :meth:`OracleSim.jv`           jacobian_vector_product!, newton_krylov.jl:504-671
* :func:`gmres`                  restarted GMRES (IterativeSolvers 0.9.4 gmres!, as called
                                 at newton_krylov.jl:448-450; iterate-level parity
                                 unpinned, only the converged solution is pinned)

Arrays of points are numpy ``(n, 3)`` C-contiguous, which is the memory layout of
Julia's column-major ``Matrix{Float64}(3, n)``.  Indices are 0-based.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpdx_oracle.so")

ORC_ERR_DUPLICATE = -6

_lib = None


def build(force=False):
    """Compile pdx_oracle.c with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)


# constitutive model codes of the oracle (and of the library): constitutive_models.jl
MODELS = {"SVK": 0, "NHP": 1, "LE": 2, "NH": 3}
ZEMS = {"Silling": 0, "Wan": 1}  # zero-energy-mode stabilisations of CMaterial


def pk1(model, F, lam, mu, G, K, nu):
    """first_piola_kirchhoff(model, F) of a 3x3 F (numpy, row-major), as a 3x3 array."""
    Fc = np.ascontiguousarray(np.asarray(F, np.float64).T.reshape(-1))  # column-major
    P = np.zeros(9)
    lib().orc_pk1(MODELS[model], _p(Fc), lam, mu, G, K, nu, _p(P))
    return P.reshape(3, 3).T


def zem_c1(Kinv, lam, mu, nu):
    """calc_zem_stiffness_tensor(C, Kinv) of the Hooke tensor (nu, lambda, mu), 3x3."""
    Kc = np.ascontiguousarray(np.asarray(Kinv, np.float64).T.reshape(-1))
    C1 = np.zeros(9)
    lib().orc_zem_c1(_p(Kc), lam, mu, nu, _p(C1))
    return C1.reshape(3, 3).T


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double
_INT = ctypes.c_int


def _declare(L):
    L.orc_find_bonds_count.restype = _I64
    L.orc_find_bonds_count.argtypes = [_I64, _P, _P, _P, _P, _INT]
    L.orc_find_bonds_fill.restype = _INT
    L.orc_find_bonds_fill.argtypes = [_I64, _P, _P, _P, _P, _P, _P, _P, _INT]
    L.orc_find_bonds_brute.restype = _I64
    L.orc_find_bonds_brute.argtypes = [_I64, _P, _P, _P, _P, _P, _P]
    L.orc_precrack_hits.restype = None
    L.orc_precrack_hits.argtypes = [_I64, _P, _P, _P, _P, _P]
    L.orc_bb_force.restype = None
    L.orc_bb_force.argtypes = [_I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _D, _D, _P, _P, _INT, _P]
    L.orc_reverse_index.restype = _INT
    L.orc_reverse_index.argtypes = [_I64, _P, _P, _P, _P, _P]
    L.orc_osb_force.restype = _INT
    L.orc_osb_force.argtypes = [_I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _D, _D,
                                _D, _P, _P, _P, _P, _P, _P, _P, _P, _INT]
    L.orc_cpd_force.restype = _INT
    L.orc_cpd_force.argtypes = [_I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _INT, _D, _D,
                                _D, _D, _D, _D, _D, _D, _D, _D, _P, _P, _P, _P, _P, _P, _P,
                                _INT]
    L.orc_cpd_force_sets.restype = _INT
    L.orc_cpd_force_sets.argtypes = [_I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _INT,
                                     _P, _P, _P, _P, _P, _P, _P, _D, _D, _D, _P, _P, _P, _P,
                                     _P, _P, _P, _INT]
    L.orc_zem_c1.restype = None
    L.orc_zem_c1.argtypes = [_P, _D, _D, _D, _P]
    L.orc_pk1.restype = None
    L.orc_pk1.argtypes = [_INT, _P, _D, _D, _D, _D, _D, _P]
    L.orc_min_timestep.restype = _D
    L.orc_min_timestep.argtypes = [_I64, _P, _P, _P, _P, _D, _D]
    L.orc_update_vel_half.restype = None
    L.orc_update_vel_half.argtypes = [_I64, _P, _P, _D, _P]
    L.orc_update_disp_and_pos.restype = None
    L.orc_update_disp_and_pos.argtypes = [_I64, _P, _D, _P, _P]
    L.orc_update_acc_and_vel.restype = None
    L.orc_update_acc_and_vel.argtypes = [_I64, _P, _P, _D, _P, _D, _P, _P]
    L.orc_bb_force_sets.restype = None
    L.orc_bb_force_sets.argtypes = [_I64] + [_P] * 14
    L.orc_osb_force_sets.restype = _INT
    L.orc_osb_force_sets.argtypes = [_I64] + [_P] * 21 + [_INT]
    L.orc_min_timestep_sets.restype = _D
    L.orc_min_timestep_sets.argtypes = [_I64, _P, _P, _P, _P, _P, _P]
    L.orc_update_acc_and_vel_sets.restype = None
    L.orc_update_acc_and_vel_sets.argtypes = [_I64, _P, _P, _P, _P, _D, _P, _P]
    L.orc_esc_factors.restype = None
    L.orc_esc_factors.argtypes = [_I64, _INT, _INT, _P, _P, _P, _P, _P, _P, _D, _P, _P, _P, _P,
                                  _P, _P, _P, _P]
    L.orc_has_nan.restype = _INT
    L.orc_has_nan.argtypes = [_I64, _P]
    L.orc_max_threads.restype = _INT
    L.orc_max_threads.argtypes = []


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def influence(kernel, delta, L):
    """The influence function omega(delta, L) of every bond (find_kernels,
    bond_system.jl:184-196), vectorised over L, as src/discretization/kernels.jl evaluates
    it: "linear" delta / L (:22), "const_one" 1 (:10), "cubic_b_spline" (:43-52) and
    "cubic_b_spline_norm" (:74-84) of xi = L / delta, with Julia's literal powers
    (xi^2 = xi xi, xi^3 = xi xi xi) and left-to-right sums."""
    L = np.asarray(L, np.float64)
    if kernel == "linear":
        return delta / L
    if kernel == "const_one":
        return np.ones_like(L)
    xi = L / delta
    x2, x3 = xi * xi, xi * xi * xi
    lo = (xi > 0) & (xi <= 0.5)
    hi = (xi > 0.5) & (xi <= 1.0)
    if kernel == "cubic_b_spline":
        a = ((2 / 3) - 4.0 * x2) + 4.0 * x3
        b = (((4 / 3) - 4.0 * xi) + 4.0 * x2) - (4 / 3) * x3
        return np.where(lo, a, np.where(hi, b, 0.0))
    if kernel == "cubic_b_spline_norm":
        C = 8.0 / (np.pi * (delta * delta * delta))
        t = 1.0 - xi
        a = C * ((1.0 - 6.0 * x2) + 6.0 * x3)
        b = (C * 2.0) * (t * t * t)
        return np.where(lo, a, np.where(hi, b, 0.0))
    raise ValueError(kernel)


class DuplicatePointError(RuntimeError):
    pass


def find_bonds(X, delta, fail_permit=None, nthreads=0, brute=False):
    """Bond list of every point: (row_ptr int64[n+1], nbr int32, L float64, bond_fp uint8).

    Neighbours of a point are listed in ascending index order (see pdx_oracle.c)."""
    X = _c(X, np.float64).reshape(-1, 3)
    n = X.shape[0]
    delta = _c(np.broadcast_to(np.asarray(delta, dtype=np.float64), (n,)), np.float64)
    fp = _c(np.ones(n, np.uint8) if fail_permit is None else np.asarray(fail_permit, np.uint8),
            np.uint8)
    counts = np.zeros(n, np.int64)
    if brute:
        total = lib().orc_find_bonds_brute(n, _p(X), _p(delta), _p(counts), None, None, None)
        row_ptr = np.zeros(n + 1, np.int64)
        np.cumsum(counts, out=row_ptr[1:])
        nbr = np.zeros(total, np.int32)
        L = np.zeros(total, np.float64)
        lib().orc_find_bonds_brute(n, _p(X), _p(delta), None, _p(row_ptr), _p(nbr), _p(L))
        bfp = (fp[np.repeat(np.arange(n), counts)] & fp[nbr]).astype(np.uint8)
        if total and np.min(L) < np.finfo(np.float64).eps:
            raise DuplicatePointError("point duplicate found")
        return row_ptr, nbr, L, bfp
    dup = np.zeros(2, np.int64)
    total = lib().orc_find_bonds_count(n, _p(X), _p(delta), _p(counts), _p(dup), nthreads)
    if total == ORC_ERR_DUPLICATE:
        raise DuplicatePointError(
            f"point duplicate found! Point #{dup[0] + 1} has a duplicate #{dup[1] + 1}")
    if total < 0:
        raise RuntimeError(f"orc_find_bonds_count failed: {total}")
    row_ptr = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=row_ptr[1:])
    nbr = np.zeros(total, np.int32)
    L = np.zeros(total, np.float64)
    bfp = np.zeros(total, np.uint8)
    st = lib().orc_find_bonds_fill(n, _p(X), _p(delta), _p(fp), _p(row_ptr), _p(nbr), _p(L),
                                   _p(bfp), nthreads)
    if st != 0:
        raise RuntimeError(f"orc_find_bonds_fill failed: {st}")
    return row_ptr, nbr, L, bfp


def precrack_hits(row_ptr, nbr, in_a, in_b):
    n = row_ptr.shape[0] - 1
    hit = np.zeros(nbr.shape[0], np.uint8)
    lib().orc_precrack_hits(n, _p(row_ptr), _p(nbr), _p(_c(in_a, np.uint8)),
                            _p(_c(in_b, np.uint8)), _p(hit))
    return hit.astype(bool)


def _sets_to_mask(n, pts):
    m = np.zeros(n, np.uint8)
    m[np.asarray(pts, dtype=np.int64)] = 1
    return m


class OracleSim:
    """State of one body on the CPU, laid out as the reference's single chunk.

    kind: "BB", "OSB" or "CPD".  params: dict with delta, rho, E, nu, G, K, lambda, mu,
    Gc, epsc, bc (as StandardPointParameters / CPointParameters), plus for CPD:
    model ("SVK" | "NHP"), Cs, maxdmg.
    precracks: list of (points_a, points_b, filter_bonds) as PointSetsPreCrack.
    """

    def __init__(self, X, vol, kind, params, fail_permit=None, precracks=(), nthreads=0,
                 bonds=None, point_set=None, correction=None):
        self.X = _c(X, np.float64).reshape(-1, 3).copy()
        n = self.X.shape[0]
        self.n = n
        self.vol = _c(np.broadcast_to(np.asarray(vol, np.float64), (n,)), np.float64).copy()
        self.kind = kind
        # several parameter sets: params is a list of dicts and point_set[i] the set of point
        # i (ParameterHandler, parameter_handler.jl:36-81); self.pt holds per-point arrays
        self.pt = None
        if point_set is not None:
            sets = [dict(q) for q in params]
            ps = np.asarray(point_set, np.int64)
            if kind not in ("BB", "OSB", "CPD"):
                raise ValueError(kind)
            self.pt = {k: np.ascontiguousarray([sets[m][k] for m in ps], np.float64)
                       for k in ("bc", "epsc", "G", "K", "rho", "lambda", "mu", "nu")}
            params = sets[0]
        self.p = dict(params)
        self.nthreads = nthreads
        fp = np.ones(n, np.uint8) if fail_permit is None else np.asarray(fail_permit, np.uint8)
        if bonds is None:
            bonds = find_bonds(self.X, self.p["delta"], fp, nthreads)
        row_ptr, nbr, L, bfp = bonds
        counts = np.diff(row_ptr)
        # precracks with filter_bonds = true delete the bonds (filter_bonds_by_crack!)
        for a, b, filt in precracks:
            if not filt:
                continue
            hit = precrack_hits(row_ptr, nbr, _sets_to_mask(n, a), _sets_to_mask(n, b))
            keep = ~hit
            owner = np.repeat(np.arange(n), counts)
            counts = np.bincount(owner[keep], minlength=n).astype(np.int64)
            row_ptr = np.zeros(n + 1, np.int64)
            np.cumsum(counts, out=row_ptr[1:])
            nbr, L, bfp = nbr[keep].copy(), L[keep].copy(), bfp[keep].copy()
        self.row_ptr, self.nbr, self.L, self.bfp = row_ptr, nbr, L, bfp
        self.n_neighbors = np.diff(row_ptr)
        nb = nbr.shape[0]
        self.active = np.ones(nb, np.uint8)
        self.bond_length = np.zeros(nb, np.float64)
        self.n_active = self.n_neighbors.astype(np.int64).copy()
        self.damage = np.zeros(n, np.float64)
        # precracks with filter_bonds = false break the bonds at t = 0 (break_bonds!)
        for a, b, filt in precracks:
            if filt:
                continue
            hit = precrack_hits(row_ptr, nbr, _sets_to_mask(n, a), _sets_to_mask(n, b))
            self.active[hit] = 0
        owner = np.repeat(np.arange(n), self.n_neighbors)
        self.n_active = np.bincount(owner, weights=self.active, minlength=n).astype(np.int64)
        with np.errstate(invalid="ignore", divide="ignore"):
            self.damage = 1.0 - self.n_active / self.n_neighbors
        # kernels (find_kernels): omega(delta_i, L) per bond for OSB and CPD (params
        # "kernel": "linear" (default), "const_one", "cubic_b_spline", "cubic_b_spline_norm")
        if kind in ("OSB", "CPD"):
            self.omega = influence(self.p.get("kernel", "linear"), self.p["delta"], L)
        else:
            self.omega = None
        # reverse bond index of the two-phase OSB/CPD loops (built once: the bond list is fixed)
        self.rev = None
        if kind in ("OSB", "CPD"):
            in_ptr = np.zeros(n + 1, np.int64)
            in_bond = np.zeros(max(nb, 1), np.int64)
            in_src = np.zeros(max(nb, 1), np.int32)
            if lib().orc_reverse_index(n, _p(row_ptr), _p(nbr), _p(in_ptr), _p(in_bond),
                                       _p(in_src)) != 0:
                raise MemoryError("orc_reverse_index")
            self.rev = (in_ptr, in_bond, in_src)
        z = lambda: np.zeros((n, 3), np.float64)
        self.x = self.X.copy()
        self.u, self.v, self.vh, self.a, self.b_int, self.b_ext = z(), z(), z(), z(), z(), z()
        # per point and component: sum of |bond contribution| of the last force evaluation,
        # the scale of the rounding error of the (cancelling) force sum
        self.b_abs = z()
        self.defgrad = np.zeros((n, 9), np.float64)
        self.cauchy = np.zeros((n, 9), np.float64)
        self.wvol = np.zeros(n, np.float64)
        self.dil = np.zeros(n, np.float64)
        # EnergySurfaceCorrection: factors from the bond states after the precracks
        # (initialize!, bond_system_corrections.jl:76-87)
        self.scf = None
        self.mfactor = None
        if correction == "energy":
            if kind not in ("BB", "OSB"):
                raise ValueError("EnergySurfaceCorrection: BB and OSB")
            if self.pt is not None:
                pp = {k: self.pt[k] for k in ("bc", "G", "K")}
                pp["lambda"] = self.pt["lambda"]
                pp["mu"] = self.pt["mu"]
            else:
                pp = {k: np.full(n, float(self.p[k])) for k in ("bc", "G", "K", "lambda", "mu")}
            self.mfactor = np.zeros((n, 3), np.float64)
            self.scf = np.zeros(nb, np.float64)
            lib().orc_esc_factors(n, 0 if kind == "BB" else 1, 1 if self.pt is not None else 0,
                                  _p(self.X), _p(self.vol), _p(self.row_ptr), _p(self.nbr),
                                  _p(self.L), _p(self.active), float(self.p["delta"]),
                                  _p(pp["lambda"]), _p(pp["mu"]), _p(pp["bc"]), _p(pp["G"]),
                                  _p(pp["K"]), _p(self.mfactor), _p(self.scf),
                                  _p(self.omega) if kind == "OSB" and
                                  self.p.get("kernel", "linear") != "linear" else None)
        elif correction is not None:
            raise ValueError(correction)

    def _rev_args(self):
        in_ptr, in_bond, in_src = self.rev
        return _p(in_ptr), _p(in_bond), _p(in_src), self.nthreads

    @staticmethod
    def _check(rc):
        if rc != 0:
            raise MemoryError("oracle force loop: out of host memory")

    # calc_force_density!(chunk, t, dt)
    def force(self, x=None):
        if x is not None:
            self.x[...] = x
        p = self.p
        L = lib()
        if self.pt is not None:
            t = self.pt
            if self.kind == "BB":
                L.orc_bb_force_sets(self.n, _p(self.x), _p(self.vol), _p(self.row_ptr),
                                    _p(self.nbr), _p(self.L), _p(self.bfp), _p(self.active),
                                    _p(self.bond_length), _p(self.n_active), _p(self.damage),
                                    _p(t["bc"]), _p(t["epsc"]), _p(self.b_int), _p(self.b_abs),
                                    _p(self.scf))
            elif self.kind == "CPD":
                model = MODELS[p.get("model", "SVK")] | (ZEMS[p.get("zem", "Silling")] << 8)
                self._check(L.orc_cpd_force_sets(self.n, _p(self.X), _p(self.x), _p(self.vol),
                                     _p(self.row_ptr), _p(self.nbr), _p(self.L), _p(self.bfp),
                                     _p(self.active), _p(self.omega), _p(self.n_active),
                                     _p(self.damage), model, _p(t["lambda"]), _p(t["mu"]),
                                     _p(t["G"]), _p(t["K"]), _p(t["nu"]), _p(t["bc"]),
                                     _p(t["epsc"]),
                                     p["delta"], p.get("Cs", 0.5), p.get("maxdmg", 0.85),
                                     _p(self.defgrad), _p(self.cauchy), _p(self.b_int),
                                     _p(self.b_abs), *self._rev_args()))
            else:
                self._check(L.orc_osb_force_sets(self.n, _p(self.X), _p(self.x), _p(self.vol),
                                     _p(self.row_ptr), _p(self.nbr), _p(self.L), _p(self.bfp),
                                     _p(self.active), _p(self.bond_length), _p(self.omega),
                                     _p(self.n_active), _p(self.damage), _p(t["G"]),
                                     _p(t["K"]), _p(t["epsc"]), _p(self.b_int), _p(self.b_abs),
                                     _p(self.scf), *self._rev_args()))
            return self.b_int
        if self.kind == "BB":
            L.orc_bb_force(self.n, _p(self.x), _p(self.vol), _p(self.row_ptr), _p(self.nbr),
                           _p(self.L), _p(self.bfp), _p(self.active), _p(self.bond_length),
                           _p(self.n_active), _p(self.damage), p["bc"], p["epsc"],
                           _p(self.b_int), _p(self.b_abs), self.nthreads, _p(self.scf))
        elif self.kind == "OSB":
            self._check(L.orc_osb_force(self.n, _p(self.X), _p(self.x), _p(self.vol), _p(self.row_ptr),
                            _p(self.nbr), _p(self.L), _p(self.bfp), _p(self.active),
                            _p(self.bond_length), _p(self.omega), _p(self.n_active),
                            _p(self.damage), p["G"], p["K"], p["epsc"], _p(self.b_int),
                            _p(self.wvol), _p(self.dil), _p(self.b_abs), _p(self.scf),
                            *self._rev_args()))
        elif self.kind == "CPD":
            model = MODELS[p.get("model", "SVK")] | (ZEMS[p.get("zem", "Silling")] << 8)
            self._check(L.orc_cpd_force(self.n, _p(self.X), _p(self.x), _p(self.vol), _p(self.row_ptr),
                            _p(self.nbr), _p(self.L), _p(self.bfp), _p(self.active),
                            _p(self.omega), _p(self.n_active), _p(self.damage), model,
                            p["lambda"], p["mu"], p["G"], p["K"], p.get("nu", 0.25), p["bc"],
                            p["delta"],
                            p.get("Cs", 0.5), p.get("maxdmg", 0.85), p["epsc"],
                            _p(self.defgrad), _p(self.cauchy), _p(self.b_int),
                            _p(self.b_abs), *self._rev_args()))
        else:
            raise ValueError(self.kind)
        return self.b_int

    def stable_dt(self, safety_factor=0.7):
        if self.pt is not None:
            dt = lib().orc_min_timestep_sets(self.n, _p(self.vol), _p(self.row_ptr),
                                             _p(self.nbr), _p(self.L), _p(self.pt["bc"]),
                                             _p(self.pt["rho"]))
            return dt * safety_factor
        dt = lib().orc_min_timestep(self.n, _p(self.vol), _p(self.row_ptr), _p(self.nbr),
                                    _p(self.L), self.p["bc"], self.p["rho"])
        return dt * safety_factor

    @staticmethod
    def _apply_bcs(field, Xref, bcs, t):
        # apply_boundary_conditions!: boundary_conditions.jl:212-226 (single-dim BCs first,
        # then the position-dependent ones), a non-finite value leaves the field as it is.
        for kind, pts, dim, f in bcs:
            pts = np.asarray(pts, np.int64)
            if kind == "sd":
                val = float(f(t))
                if np.isfinite(val):
                    field[pts, dim] = val
        for kind, pts, dim, f in bcs:
            pts = np.asarray(pts, np.int64)
            if kind == "pd":
                for i in pts:
                    val = float(f(Xref[i], t))
                    if np.isfinite(val):
                        field[i, dim] = val
            elif kind == "pdv":
                # the same, for an f that takes all the points' positions at once (k x 3)
                val = np.asarray(f(Xref[pts], t), np.float64)
                ok = np.isfinite(val)
                field[pts[ok], dim] = val[ok]

    def vv_run(self, n_steps, dt, bcs=(), first_step=1, check_nans=True, fd_bcs=()):
        """solve!(dh, VelocityVerlet): verlet_timestep! for n = first_step .. (velocity_verlet.jl:181-197).

        bcs: list of ("sd", points, dim, f(t)) and ("pd", points, dim, f(X, t)) acting on
        velocity_half (velocity_bc!, boundary_conditions.jl:422-435); fd_bcs the same on
        b_ext (forcedensity_bc!, :559-575), both applied by apply_boundary_conditions! after
        the half step; ("pdv", points, dim, f(Xs, t)) is "pd" with f vectorised over the
        points' reference positions Xs (k x 3)."""
        L = lib()
        ndof = 3 * self.n
        dth = 0.5 * dt
        for n in range(first_step, first_step + n_steps):
            t = n * dt
            L.orc_update_vel_half(ndof, _p(self.v), _p(self.a), dth, _p(self.vh))
            self._apply_bcs(self.vh, self.X, bcs, t)
            self._apply_bcs(self.b_ext, self.X, fd_bcs, t)
            L.orc_update_disp_and_pos(ndof, _p(self.vh), dt, _p(self.u), _p(self.x))
            self.force()
            if check_nans and L.orc_has_nan(ndof, _p(self.b_int)):
                raise FloatingPointError(f"NaN in b_int at t={t}, step {n}")
            if self.pt is not None:
                L.orc_update_acc_and_vel_sets(self.n, _p(self.b_int), _p(self.b_ext),
                                              _p(self.pt["rho"]), _p(self.vh), dth, _p(self.a),
                                              _p(self.v))
            else:
                L.orc_update_acc_and_vel(ndof, _p(self.b_int), _p(self.b_ext), self.p["rho"],
                                         _p(self.vh), dth, _p(self.a), _p(self.v))
        return self

    # ---------------------------------------------------------------- dynamic relaxation
    def dr_density(self, dt, damping_factor):
        """calc_density_matrix! for BondSystem: dynamic_relaxation.jl:166-176 (per point)."""
        bc = self.pt["bc"] if self.pt is not None else self.p["bc"]  # get_params(paramsetup, i)
        k = ((5.0 * np.pi) * (self.p["delta"] * self.p["delta"])) * bc
        return ((((damping_factor * 1.0) / 4.0) * (dt * dt)) * self.n_neighbors.astype(np.float64)) * k

    def dr_damping(self, dt, density, vh_old, bint_old):
        """calc_damping: dynamic_relaxation.jl:262-323, the sums taken exactly (math.fsum)."""
        import math
        u = self.u.reshape(-1)
        vh = self.vh.reshape(-1)
        rho = np.repeat(density, 3)
        m = vh != 0.0
        db = self.b_int.reshape(-1)[m] - bint_old.reshape(-1)[m]
        temp = (dt * rho[m]) * vh_old.reshape(-1)[m]
        cn1 = math.fsum(((-(u[m] * u[m])) * db) / temp)
        cn2 = math.fsum(u * u)
        cn = 0.0
        if cn2 != 0.0 and cn1 / cn2 > 0.0:
            cn = 2.0 * math.sqrt(cn1 / cn2)
        if cn > 2.0:
            cn = 1.9
        return cn

    def dr_run(self, n_steps, dt, damping_factor=1.0, bcs=(), fd_bcs=(), state=None):
        """solve!(dh, DynamicRelaxation): relaxation_timestep! for n = 1 .. n_steps
        (dynamic_relaxation.jl:219-350).  bcs act on velocity_half, fd_bcs on b_ext, both as
        ("sd", points, dim, f(t)) or ("pd", points, dim, f(X, t)).  Returns the state dict
        (density, vh_old, b_int_old) to continue with."""
        L = lib()
        if state is None:
            state = {"density": self.dr_density(dt, damping_factor),
                     "vh_old": np.zeros((self.n, 3)), "bint_old": np.zeros((self.n, 3)), "n": 0}
        rho = state["density"][:, None]
        for _ in range(n_steps):
            n = state["n"] + 1
            t = n * dt
            self._apply_bcs(self.vh, self.X, bcs, t)
            self._apply_bcs(self.b_ext, self.X, fd_bcs, t)
            L.orc_update_disp_and_pos(3 * self.n, _p(self.vh), dt, _p(self.u), _p(self.x))
            self.force()
            if n == 1:
                vn = ((0.5 * dt) * (self.b_int + self.b_ext)) / rho
            else:
                cn = self.dr_damping(dt, state["density"], state["vh_old"], state["bint_old"])
                a = (self.b_int + self.b_ext) / rho
                vn = ((2.0 - cn * dt) * self.vh + (2.0 * dt) * a) / (2.0 + cn * dt)
            self.vh[...] = vn
            self.v[...] = 0.5 * (state["vh_old"] + vn)
            state["vh_old"] = vn.copy()
            state["bint_old"] = self.b_int.copy()
            state["n"] = n
        return state

    # ---------------------------------------------------------------- Newton-Krylov
    def residual(self, constrained):
        """calc_residual!: newton_krylov.jl:371-387."""
        r = (self.b_int + self.b_ext) * self.vol[:, None]
        r = r.reshape(-1)
        r[constrained] = 0.0
        return r

    def jv(self, v, constrained, scale=1.0):
        """jacobian_vector_product!: newton_krylov.jl:504-532 and :537-671."""
        v = v.reshape(self.n, 3)
        u_norm = np.sqrt(np.sum(self.u.reshape(-1) ** 2))
        v_norm = np.sqrt(np.sum(v.reshape(-1) ** 2))
        if v_norm < np.finfo(np.float64).eps:
            v_norm = 1.0
        eps = scale * np.cbrt(np.finfo(np.float64).eps) * (u_norm + 1.0) / v_norm
        u0 = self.u.copy()
        b0 = self.b_int.copy()
        self.u[...] = u0 + eps * v
        self.x[...] = self.X + self.u
        fp = self.force() * self.vol[:, None]
        fp = fp.copy()
        self.u[...] = u0 - eps * v
        self.x[...] = self.X + self.u
        fm = self.force() * self.vol[:, None]
        jv = ((fp - fm) / (2 * eps)).reshape(-1)
        jv[constrained] = v.reshape(-1)[constrained]
        self.u[...] = u0
        self.x[...] = self.X + self.u
        self.b_int[...] = b0
        return jv

    def nk_step(self, n, n_steps, disp_bcs, maxiter=100, tol=1e-4, rel_tol=None,
                gmres_maxiter=200, gmres_reltol=1e-4, gmres_abstol=1e-8, gmres_restart=100,
                scale=1.0, log=None):
        """newton_krylov_step!: newton_krylov.jl:298-332.

        disp_bcs: list of (points, dim, f(X)) (displacement_bc!, PosSingleDimBC).
        tol is absolute as in the reference; rel_tol (optional) stops at
        ||r|| < rel_tol * ||r_first|| of this load step instead."""
        beta = n / n_steps
        constrained = self.constrained_dofs(disp_bcs)
        # apply_incr_boundary_conditions! + update_position! (constrained dofs only)
        for pts, dim, f in disp_bcs:
            for i in np.asarray(pts, np.int64):
                val = beta * float(f(self.X[i]))
                if np.isfinite(val):
                    self.u[i, dim] = val
        xf, Xf, uf = self.x.reshape(-1), self.X.reshape(-1), self.u.reshape(-1)
        xf[constrained] = Xf[constrained] + uf[constrained]
        free = np.ones(3 * self.n, bool)
        free[constrained] = False
        r0 = None
        n_jv = 0
        for it in range(1, maxiter + 1):
            self.force()
            r = self.residual(constrained)
            rn = float(np.sqrt(np.sum(r * r)))
            if r0 is None:
                r0 = rn
            if log is not None:
                log.append((n, it, rn))
            thr = tol if rel_tol is None else rel_tol * r0
            if rn < thr:
                return {"iters": it, "residual": rn, "n_jv": n_jv}
            if it == maxiter:
                raise RuntimeError("Newton-Krylov solver did not converge after max iterations!")
            def A(v):
                nonlocal n_jv
                n_jv += 1
                return self.jv(v, constrained, scale)
            du = gmres(A, -r, maxiter=gmres_maxiter, reltol=gmres_reltol, abstol=gmres_abstol,
                       restart=gmres_restart)
            uf = self.u.reshape(-1)
            xf = self.x.reshape(-1)
            uf[free] += du[free]
            xf[free] += du[free]

    def constrained_dofs(self, disp_bcs):
        dofs = []
        for pts, dim, f in disp_bcs:
            dofs.extend(3 * int(i) + dim for i in pts)
        return np.array(sorted(set(dofs)), np.int64)


def gmres(A, b, maxiter=200, reltol=1e-4, abstol=1e-8, restart=100):
    """Restarted GMRES with modified Gram-Schmidt and Givens rotations, x0 = 0.

    Stops when the residual estimate drops below max(reltol * ||b||, abstol) or after
    maxiter matrix-vector products (IterativeSolvers' stopping rule)."""
    nrm = np.linalg.norm
    x = np.zeros_like(b)
    bnorm = nrm(b)
    thr = max(reltol * bnorm, abstol)
    it = 0
    r = b.copy()
    beta = nrm(r)
    while it < maxiter and beta > thr:
        m = restart
        V = np.zeros((m + 1, b.shape[0]))
        H = np.zeros((m + 1, m))
        cs = np.zeros(m)
        sn = np.zeros(m)
        g = np.zeros(m + 1)
        g[0] = beta
        V[0] = r / beta
        k = 0
        for k in range(m):
            w = A(V[k])
            it += 1
            for j in range(k + 1):
                H[j, k] = np.dot(V[j], w)
                w = w - H[j, k] * V[j]
            H[k + 1, k] = nrm(w)
            if H[k + 1, k] != 0:
                V[k + 1] = w / H[k + 1, k]
            for j in range(k):
                t = cs[j] * H[j, k] + sn[j] * H[j + 1, k]
                H[j + 1, k] = -sn[j] * H[j, k] + cs[j] * H[j + 1, k]
                H[j, k] = t
            den = np.hypot(H[k, k], H[k + 1, k])
            cs[k] = H[k, k] / den if den != 0 else 1.0
            sn[k] = H[k + 1, k] / den if den != 0 else 0.0
            H[k, k] = den
            H[k + 1, k] = 0.0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            if abs(g[k + 1]) <= thr or it >= maxiter:
                k += 1
                break
        else:
            k = m
        y = np.linalg.solve(np.triu(H[:k, :k]), g[:k]) if k > 0 else np.zeros(0)
        x = x + V[:k].T @ y
        r = b - A(x)
        it += 1
        beta = nrm(r)
    return x
