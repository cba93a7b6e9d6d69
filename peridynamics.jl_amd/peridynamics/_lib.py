"""ctypes binding of libpdx.so (the C ABI declared in include/pdx.h).

The product path has no fallback: if libpdx.so is missing or no GPU is visible, every
call raises.  This is the binding a Julia ``ccall`` shim mirrors one to one."""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PDX_LIB") or os.path.join(os.path.dirname(PKG_DIR), "libpdx.so")

PDX_OK = 0
PDX_ERR_ARG = -1
PDX_ERR_HIP = -2
PDX_ERR_OOM = -3
PDX_ERR_NAN = -4
PDX_ERR_STATE = -5
PDX_ERR_DUPLICATE_POINT = -6
PDX_ERR_NOT_CONVERGED = -7
PDX_ERR_UNSUPPORTED = -8

PDX_FP64 = 0
PDX_MIXED = 1

PDX_BB, PDX_OSB, PDX_CPD = 0, 1, 2
PDX_SVK, PDX_NHP, PDX_LINEAR_ELASTIC, PDX_NEO_HOOKE = 0, 1, 2, 3
PDX_ZEM_SILLING, PDX_ZEM_WAN = 0, 1

FIELDS = {
    "position": (0, 3),
    "displacement": (1, 3),
    "velocity": (2, 3),
    "velocity_half": (3, 3),
    "acceleration": (4, 3),
    "b_int": (5, 3),
    "b_ext": (6, 3),
    "damage": (7, 1),
    "n_active_bonds": (8, 1),
    "defgrad": (9, 9),
    "cauchy_stress": (10, 9),
    "ref_position": (11, 3),
    "volume": (12, 1),
    "velocity_half_old": (13, 3),
    "b_int_old": (14, 3),
    "density_matrix": (15, 1),
}


class PdxError(RuntimeError):
    """A libpdx call failed."""

    def __init__(self, status, msg):
        super().__init__(f"[pdx status {status}] {msg}")
        self.status = status


class ArgumentError(PdxError, ValueError):
    """Invalid argument (the reference's ArgumentError)."""


class NaNError(PdxError, FloatingPointError):
    """NaN in the force density (the reference's NaNError(t, step), errors.jl:98-114)."""

    def __init__(self, status, msg, step=None, t=None):
        super().__init__(status, msg)
        self.step = step
        self.t = t


class DuplicatePointError(PdxError):
    """check_point_duplicates (bond_system.jl:125-132)."""


class NotConvergedError(PdxError):
    """Newton-Krylov did not converge (newton_krylov.jl:318-324)."""


class PdxParams(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in
                ("delta", "rho", "E", "nu", "G", "K", "lambda_", "mu", "Gc", "epsc", "bc")] + [
        ("model", ctypes.c_int32), ("kernel", ctypes.c_int32), ("zem_cs", ctypes.c_double),
        ("maxdmg", ctypes.c_double), ("zem", ctypes.c_int32), ("reserved_", ctypes.c_int32)]


class PdxNkOpts(ctypes.Structure):
    _fields_ = [("maxiter", ctypes.c_int32), ("gmres_maxiter", ctypes.c_int32),
                ("gmres_restart", ctypes.c_int32), ("precond", ctypes.c_int32),
                ("tol", ctypes.c_double), ("rel_tol", ctypes.c_double),
                ("gmres_reltol", ctypes.c_double), ("gmres_abstol", ctypes.c_double),
                ("perturbation_scale", ctypes.c_double)]


class PdxNkStats(ctypes.Structure):
    _fields_ = [("newton_iters", ctypes.c_int32), ("krylov_iters", ctypes.c_int32),
                ("jv_evals", ctypes.c_int32), ("precond_fallbacks", ctypes.c_int32),
                ("residual_first", ctypes.c_double), ("residual_final", ctypes.c_double),
                ("seconds", ctypes.c_double)]


_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_D = ctypes.c_double
_ST = ctypes.c_int32

# name -> (restype, argtypes): every entry point of include/pdx.h
SIGNATURES = {
    "pdx_version": (ctypes.c_char_p, []),
    "pdx_create": (_ST, [ctypes.POINTER(_P), _I32, _I32]),
    "pdx_destroy": (None, [_P]),
    "pdx_last_error": (ctypes.c_char_p, [_P]),
    "pdx_synchronize": (_ST, [_P]),
    "pdx_set_points": (_ST, [_P, _I64, _P, _P, _P]),
    "pdx_set_material": (_ST, [_P, _I32, ctypes.POINTER(PdxParams)]),
    "pdx_set_material_sets": (_ST, [_P, _I32, _I32, ctypes.POINTER(PdxParams), _P]),
    "pdx_set_surface_correction": (_ST, [_P, _I32]),
    "pdx_export_begin": (_ST, [_P, _I32, _P, _I64]),
    "pdx_export_wait": (_ST, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I32),
                              ctypes.POINTER(ctypes.c_void_p)]),
    "pdx_get_surface_correction": (_ST, [_P, _P, _P]),
    "pdx_add_precrack": (_ST, [_P, _I64, _P, _I64, _P, _I32]),
    "pdx_build_bonds": (_ST, [_P, ctypes.POINTER(_I64)]),
    "pdx_get_bonds": (_ST, [_P, _P, _P, _P, _P, _P]),
    "pdx_set_field": (_ST, [_P, _I32, _P]),
    "pdx_get_field": (_ST, [_P, _I32, _P]),
    "pdx_stable_timestep": (_ST, [_P, _D, ctypes.POINTER(_D)]),
    "pdx_force_density": (_ST, [_P, _D, _D]),
    "pdx_set_velocity_bcs": (_ST, [_P, _I64, _P, _P, _I64]),
    "pdx_vv_run": (_ST, [_P, _I64, _I64, _D, _P, ctypes.POINTER(_I64)]),
    "pdx_vv_run_noalloc": (_ST, [_P, _I64, _I64, _D, _P, ctypes.POINTER(_I64)]),
    "pdx_set_force_density_bcs": (_ST, [_P, _I64, _P, _P]),
    "pdx_dr_init": (_ST, [_P, _D, _D]),
    "pdx_dr_run": (_ST, [_P, _I64, _I64, _D, _P, ctypes.POINTER(_I64)]),
    "pdx_dr_damping": (_ST, [_P, _D, ctypes.POINTER(_D)]),
    "pdx_set_displacement_bcs": (_ST, [_P, _I64, _P, _P, _P]),
    "pdx_jv": (_ST, [_P, _P, _P, _D]),
    "pdx_residual": (_ST, [_P, _P, ctypes.POINTER(_D)]),
    "pdx_nk_coarse_info": (_ST, [_P, ctypes.POINTER(_I32), ctypes.POINTER(_I64),
                                 ctypes.POINTER(_I32)]),
    "pdx_nk_step": (_ST, [_P, _I64, _I64, ctypes.POINTER(PdxNkOpts),
                          ctypes.POINTER(PdxNkStats)]),
    "pdx_gnn_collect": (_ST, [_P, _I32, _I64]),
    "pdx_gnn_pairs": (_ST, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    "pdx_gnn_train": (_ST, [_P, _I32, _I32, _D, ctypes.c_uint64, ctypes.POINTER(_D)]),
    "pdx_gnn_gate": (_ST, [_P, _D, ctypes.POINTER(_D), ctypes.POINTER(_I32)]),
    "pdx_gnn_n_params": (_I64, []),
    "pdx_gnn_params": (_ST, [_P, _P, _P, ctypes.POINTER(_D)]),
    "pdx_gnn_graph": (_ST, [_P, ctypes.POINTER(_I64), _P, _P, _P, _P, _P, _P]),
    "pdx_gnn_apply_vec": (_ST, [_P, _P, _P]),
    "pdx_gnn_field_values": (_ST, [_P, _P]),
    "pdx_gnn_loss_grad": (_ST, [_P, _I32, _P, _P, ctypes.POINTER(_D), _P]),
    "pdx_set_profiling": (_ST, [_P, _I32]),
    "pdx_kernel_time": (_ST, [_P, ctypes.c_char_p, ctypes.POINTER(_D), ctypes.POINTER(_I64)]),
    "pdx_info": (_ST, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    "pdx_layout": (_ST, [_P, ctypes.POINTER(_I32), ctypes.POINTER(_I64), ctypes.POINTER(_I64),
                         ctypes.POINTER(_I32)]),
}

_lib = None


def load():
    """Load libpdx.so (built in-tree by __graft_entry__.build()).  Raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"libpdx.so not found at {LIB_PATH}; run __graft_entry__.build() first")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def check(ctx, status, what=""):
    if status == PDX_OK:
        return
    msg = load().pdx_last_error(ctx)
    msg = msg.decode() if msg else ""
    if what:
        msg = f"{what}: {msg}"
    if status == PDX_ERR_ARG:
        raise ArgumentError(status, msg)
    if status == PDX_ERR_NAN:
        raise NaNError(status, msg)
    if status == PDX_ERR_DUPLICATE_POINT:
        raise DuplicatePointError(status, msg)
    if status == PDX_ERR_NOT_CONVERGED:
        raise NotConvergedError(status, msg)
    raise PdxError(status, msg)


class Context:
    """Owner of one pdx_ctx (one body on one GPU)."""

    def __init__(self, device=0, precision=PDX_FP64):
        self.lib = load()
        self.h = ctypes.c_void_p()
        st = self.lib.pdx_create(ctypes.byref(self.h), int(device), int(precision))
        if st != PDX_OK:
            raise PdxError(st, f"pdx_create(device={device}) failed (is a GPU visible?)")
        self.precision = precision

    def close(self):
        if self.h:
            self.lib.pdx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def call(self, name, *args):
        st = getattr(self.lib, name)(self.h, *args)
        check(self.h, st, name)
        return st

    # convenience wrappers ------------------------------------------------------------
    def set_points(self, X, vol, fail_permit=None):
        X = np.ascontiguousarray(X, np.float64).reshape(-1, 3)
        n = X.shape[0]
        vol = np.ascontiguousarray(np.broadcast_to(np.asarray(vol, np.float64), (n,)))
        fp = None if fail_permit is None else np.ascontiguousarray(fail_permit, np.uint8)
        self.call("pdx_set_points", n, ptr(X), ptr(vol), ptr(fp))
        self.n = n

    def set_material(self, kind, params):
        self.call("pdx_set_material", int(kind), ctypes.byref(params))

    def set_material_sets(self, kind, sets, point_set):
        """Several parameter sets: `sets` a list of PdxParams, `point_set` the 0-based set
        of every point (the reference's params_map minus one)."""
        arr = (PdxParams * len(sets))(*sets)
        ps = np.ascontiguousarray(point_set, np.int32)
        if ps.shape[0] != self.n:
            raise ArgumentError(-1, "point_set needs one entry per point")
        self.call("pdx_set_material_sets", int(kind), len(sets), arr, ptr(ps))

    def set_surface_correction(self, kind):
        self.call("pdx_set_surface_correction", int(kind))

    def get_surface_correction(self, n_bonds):
        """(scfactor per bond in the caller's bond order, mfactor (n, 3))."""
        sc = np.empty(n_bonds, np.float64)
        mf = np.empty((self.n, 3), np.float64)
        self.call("pdx_get_surface_correction", ptr(sc), ptr(mf))
        return sc, mf

    def add_precrack(self, a, b, filter_bonds):
        a = np.ascontiguousarray(a, np.int32)
        b = np.ascontiguousarray(b, np.int32)
        self.call("pdx_add_precrack", a.shape[0], ptr(a), b.shape[0], ptr(b), int(filter_bonds))

    def build_bonds(self):
        nb = _I64()
        self.call("pdx_build_bonds", ctypes.byref(nb))
        self.nb = nb.value
        return nb.value

    def get_bonds(self):
        n, nb = self.n, self.nb
        row_ptr = np.zeros(n + 1, np.int64)
        nbr = np.zeros(nb, np.int32)
        L = np.zeros(nb, np.float64)
        fp = np.zeros(nb, np.uint8)
        act = np.zeros(nb, np.uint8)
        self.call("pdx_get_bonds", ptr(row_ptr), ptr(nbr), ptr(L), ptr(fp), ptr(act))
        return row_ptr, nbr, L, fp, act

    def get_field(self, name):
        fid, comps = FIELDS[name]
        out = np.zeros((self.n, comps), np.float64)
        self.call("pdx_get_field", fid, ptr(out))
        return out if comps > 1 else out[:, 0]

    def export_begin(self, names, tag):
        """Snapshot of the named fields, copied to pinned host memory while the device
        goes on (pdx_export_begin)."""
        ids = np.ascontiguousarray([FIELDS[f][0] for f in names], np.int32)
        self.call("pdx_export_begin", len(ids), ptr(ids), int(tag))
        if not hasattr(self, "_xnames"):
            self._xnames = []
        self._xnames.append(list(names))

    def export_wait(self):
        """(tag, slot, {name: array}) of the oldest pending export.  The arrays view the
        library's pinned memory of that slot: they stay valid until the next export_begin
        that fills the same slot (the slots alternate)."""
        ptrs = (ctypes.c_void_p * 16)()
        tag, slot = _I64(), _I32()
        self.call("pdx_export_wait", ctypes.byref(tag), ctypes.byref(slot), ptrs)
        names = self._xnames.pop(0)
        out = {}
        for k, f in enumerate(names):
            comps = FIELDS[f][1]
            buf = (ctypes.c_double * (self.n * comps)).from_address(ptrs[k])
            a = np.ctypeslib.as_array(buf)
            out[f] = a.reshape(self.n, comps) if comps > 1 else a
        return tag.value, slot.value, out

    def set_field(self, name, values):
        fid, comps = FIELDS[name]
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(values, np.float64),
                                                 (self.n, comps) if comps > 1 else (self.n,)))
        self.call("pdx_set_field", fid, ptr(a))

    def stable_timestep(self, safety_factor=0.7):
        dt = _D()
        self.call("pdx_stable_timestep", float(safety_factor), ctypes.byref(dt))
        return dt.value

    def force_density(self, t=0.0, dt=0.0):
        self.call("pdx_force_density", float(t), float(dt))

    def set_velocity_bcs(self, points, src, n_src):
        points = np.ascontiguousarray(points, np.int32)
        src = np.ascontiguousarray(src, np.int32).reshape(-1)
        self.call("pdx_set_velocity_bcs", points.shape[0], ptr(points), ptr(src), int(n_src))

    def set_force_density_bcs(self, points, src):
        points = np.ascontiguousarray(points, np.int32)
        src = np.ascontiguousarray(src, np.int32).reshape(-1)
        self.call("pdx_set_force_density_bcs", points.shape[0], ptr(points), ptr(src))

    def dr_init(self, dt, damping_factor):
        self.call("pdx_dr_init", float(dt), float(damping_factor))

    def dr_run(self, first_step, n_steps, dt, bc_values=None):
        bv = None if bc_values is None else np.ascontiguousarray(bc_values, np.float64)
        nan_step = _I64(-1)
        st = self.lib.pdx_dr_run(self.h, int(first_step), int(n_steps), float(dt), ptr(bv),
                                 ctypes.byref(nan_step))
        if st == PDX_ERR_NAN:
            msg = self.lib.pdx_last_error(self.h).decode()
            raise NaNError(st, msg, step=nan_step.value, t=nan_step.value * dt)
        check(self.h, st, "pdx_dr_run")

    def dr_damping(self, dt):
        cn = _D()
        self.call("pdx_dr_damping", float(dt), ctypes.byref(cn))
        return cn.value

    def vv_run(self, first_step, n_steps, dt, bc_values=None, noalloc=False):
        """pdx_vv_run, or with noalloc=True pdx_vv_run_noalloc (the BC value table staged in
        the library's fixed device array, no device allocation)."""
        bv = None if bc_values is None else np.ascontiguousarray(bc_values, np.float64)
        nan_step = _I64(-1)
        name = "pdx_vv_run_noalloc" if noalloc else "pdx_vv_run"
        st = getattr(self.lib, name)(self.h, int(first_step), int(n_steps), float(dt), ptr(bv),
                                     ctypes.byref(nan_step))
        if st == PDX_ERR_NAN:
            msg = self.lib.pdx_last_error(self.h).decode()
            raise NaNError(st, msg, step=nan_step.value, t=nan_step.value * dt)
        check(self.h, st, name)

    def set_displacement_bcs(self, points, dims, values):
        points = np.ascontiguousarray(points, np.int32)
        dims = np.ascontiguousarray(dims, np.int32)
        values = np.ascontiguousarray(values, np.float64)
        self.call("pdx_set_displacement_bcs", points.shape[0], ptr(points), ptr(dims), ptr(values))

    def residual(self):
        r = np.zeros((self.n, 3), np.float64)
        nrm = _D()
        self.call("pdx_residual", ptr(r), ctypes.byref(nrm))
        return r, nrm.value

    def jv(self, v, scale=1.0):
        v = np.ascontiguousarray(v, np.float64).reshape(self.n, 3)
        out = np.zeros((self.n, 3), np.float64)
        self.call("pdx_jv", ptr(v), ptr(out), float(scale))
        return out

    def nk_step(self, n, n_steps, opts):
        st = PdxNkStats()
        status = self.lib.pdx_nk_step(self.h, int(n), int(n_steps), ctypes.byref(opts),
                                      ctypes.byref(st))
        if status == PDX_ERR_NOT_CONVERGED:
            msg = self.lib.pdx_last_error(self.h).decode()
            err = NotConvergedError(status, msg)
            err.stats = st
            raise err
        check(self.h, status, "pdx_nk_step")
        return st

    # ---- learned preconditioner (internal point order where noted)
    def gnn_collect(self, on=True, max_krylov_pairs=2048):
        self.call("pdx_gnn_collect", 1 if on else 0, int(max_krylov_pairs))

    def gnn_pairs(self):
        a, b = _I64(), _I64()
        self.call("pdx_gnn_pairs", ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def gnn_gate(self, threshold=float("nan")):
        """(held-out loss, network applied) of the trained network; a threshold sets the
        gate (pdx_gnn_gate)."""
        vl, use = _D(), _I32()
        self.call("pdx_gnn_gate", float(threshold), ctypes.byref(vl), ctypes.byref(use))
        return vl.value, bool(use.value)

    def gnn_train(self, steps, batch=4, lr=1e-3, seed=0):
        loss = _D()
        self.call("pdx_gnn_train", int(steps), int(batch), float(lr), ctypes.c_uint64(seed),
                  ctypes.byref(loss))
        return loss.value

    def gnn_n_params(self):
        return int(self.lib.pdx_gnn_n_params())

    def gnn_get_params(self):
        th = np.zeros(self.gnn_n_params(), np.float32)
        c = _D()
        self.call("pdx_gnn_params", ptr(th), None, ctypes.byref(c))
        return th, c.value

    def gnn_set_params(self, theta, out_scale):
        th = np.ascontiguousarray(theta, np.float32)
        if th.shape != (self.gnn_n_params(),):
            raise ValueError("parameter vector has the wrong length")
        c = _D(out_scale)
        self.call("pdx_gnn_params", None, ptr(th), ctypes.byref(c))

    def gnn_graph(self):
        """(perm, near_ptr, near, features (n, 16), dinv, cmask), internal order."""
        nn = _I64()
        self.call("pdx_gnn_graph", ctypes.byref(nn), None, None, None, None, None, None)
        n = self.n
        perm = np.zeros(n, np.int64)
        near_ptr = np.zeros(n + 1, np.int64)
        near = np.zeros(max(nn.value, 1), np.int32)
        stat = np.zeros((n, 16), np.float32)
        dinv = np.zeros((n, 9), np.float64)
        cmask = np.zeros((n, 3), np.uint8)
        self.call("pdx_gnn_graph", ctypes.byref(nn), ptr(perm), ptr(near_ptr), ptr(near),
                  ptr(stat), ptr(dinv), ptr(cmask))
        return perm, near_ptr, near[:nn.value], stat, dinv, cmask.astype(bool)

    def gnn_apply(self, y):
        y = np.ascontiguousarray(y, np.float64).reshape(self.n, 3)
        z = np.zeros_like(y)
        self.call("pdx_gnn_apply_vec", ptr(y), ptr(z))
        return z

    def coarse_info(self):
        """(points per aggregate, aggregates, coarse fields of the last coarse space)."""
        a, na, f = _I32(), _I64(), _I32()
        self.call("pdx_nk_coarse_info", ctypes.byref(a), ctypes.byref(na), ctypes.byref(f))
        return a.value, na.value, f.value

    def gnn_fields(self):
        """The network's 8 coarse fields per point, (n, 8, 3) float32, internal order."""
        f = np.zeros((self.n, 8, 3), np.float32)
        self.call("pdx_gnn_field_values", ptr(f))
        return f

    def gnn_loss_grad(self, E):
        """Loss and gradient over the error samples E (s, n, 3), internal order."""
        E = np.ascontiguousarray(E, np.float64).reshape(-1, self.n, 3)
        loss = _D()
        g = np.zeros(self.gnn_n_params(), np.float32)
        self.call("pdx_gnn_loss_grad", E.shape[0], ptr(E), None, ctypes.byref(loss), ptr(g))
        return loss.value, g

    def set_profiling(self, on=True):
        self.call("pdx_set_profiling", 1 if on else 0)

    def kernel_time(self, name):
        ms = _D()
        cnt = _I64()
        self.call("pdx_kernel_time", name.encode(), ctypes.byref(ms), ctypes.byref(cnt))
        return ms.value, cnt.value

    def synchronize(self):
        self.call("pdx_synchronize")

    def info(self):
        a, b, c = _I64(), _I64(), _I64()
        self.call("pdx_info", ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def layout(self):
        """{"lds_mode", "bond_slots", "halo_entries", "halo_max"} of the device bond lists."""
        m, s, h = _I32(), _I64(), _I64()
        hm = _I32()
        self.call("pdx_layout", ctypes.byref(m), ctypes.byref(s), ctypes.byref(h), ctypes.byref(hm))
        return {"lds_mode": bool(m.value), "bond_slots": s.value, "halo_entries": h.value,
                "halo_max": hm.value}
