"""Body, point sets, material!, failure permissions, precracks and boundary conditions.

Mirrors the user-facing set-up of the reference (src/discretization/body.jl:83-125,
point_sets.jl:73-104, predefined_cracks.jl:106-112, src/physics/material_parameters.jl:
64-92, fracture.jl:38-96, boundary_conditions.jl:257-435, 675-688).  Julia's trailing `!`
is dropped (material! -> material), point indices are 0-based, dimensions are given as
1, 2, 3 or "x", "y", "z" as in the reference."""
import inspect

import numpy as np

from ._lib import ArgumentError
from .materials import get_point_params


def _arg_error(msg):
    return ArgumentError(-1, msg)


_SYMBOL_TO_DIM = {"x": 0, "y": 1, "z": 2}


def get_dim(dim):
    """get_dim: boundary_conditions.jl:257-268 (returns 0-based)."""
    if isinstance(dim, str):
        if dim not in _SYMBOL_TO_DIM:
            raise _arg_error(f"unknown dimension {dim!r}")
        return _SYMBOL_TO_DIM[dim]
    if int(dim) not in (1, 2, 3):
        raise _arg_error(f"unknown dimension {dim!r}")
    return int(dim) - 1


def condition_function_type(f):
    """check_boundary_condition_function: boundary_conditions.jl:336-356 (argument names)."""
    try:
        names = [p.name for p in inspect.signature(f).parameters.values()]
    except (TypeError, ValueError):
        names = []
    if names == ["t"]:
        return "sdbc"
    if names == ["p"]:
        return "psdbc"
    if names == ["p", "t"]:
        return "pdsdbc"
    raise _arg_error("wrong arguments or argument names for condition function!\n"
                     "Boundary conditions support only three type of functions:\n"
                     "  `f(p, t)`\n  `f(t)`\n  `f(p)`")


class Body:
    """Body(mat, position, volume): position is 3 x n (Julia layout), volume has n entries."""

    def __init__(self, mat, position, volume):
        pos = np.asarray(position, dtype=np.float64)
        if pos.ndim != 2 or pos.shape[0] != 3:
            raise _arg_error("position must be a 3 x n array")
        n = pos.shape[1]
        vol = np.asarray(volume, dtype=np.float64).reshape(-1)
        if vol.shape[0] != n:
            raise _arg_error("number of points in volume and position have to match!")
        self.mat = mat
        self.n_points = n
        self.position = np.ascontiguousarray(pos)          # 3 x n
        self.volume = vol.copy()
        self.fail_permit = np.ones(n, dtype=bool)
        self.point_sets = {"all_points": np.arange(n, dtype=np.int64)}
        self.point_params = []
        self.params_map = np.zeros(n, dtype=np.int64)      # 0 = no material yet
        self.single_dim_bcs = []                            # (field, set, dim, f(t))
        self.posdep_single_dim_bcs = []                     # (field, set, dim, f(p, t))
        self.pos_single_dim_bcs = []                        # (field, set, dim, f(p))
        self.point_sets_precracks = []                      # (set_a, set_b, filter_bonds)
        self.single_dim_ics = []                            # (field, set, dim, value)
        self.posdep_single_dim_ics = []                     # (field, set, dim, f(p))

    @property
    def xyz(self):
        """Positions as an (n, 3) C-contiguous array (the memory of Julia's 3 x n)."""
        return np.ascontiguousarray(self.position.T)

    def __repr__(self):
        return f"{self.n_points}-point Body{{{self.mat!r}}}"


def n_points(body):
    return body.n_points


def _check_set(body, name):
    if name not in body.point_sets:
        raise _arg_error(f"there is no point set with name `{name}`!")


def point_set(body, name, points):
    """point_set!(body, name, points) / point_set!(f, body, name): point_sets.jl:73-104.

    points: 0-based indices, a boolean mask of length n, or f(p) evaluated on reference
    positions (p[0], p[1], p[2])."""
    if name in body.point_sets:
        raise _arg_error(f"there is already a point set with name `{name}`!")
    if callable(points):
        X = body.position
        try:
            mask = np.asarray(points(X), dtype=bool)
            if mask.shape != (body.n_points,):
                raise ValueError
        except Exception:
            mask = np.array([bool(points(X[:, i])) for i in range(body.n_points)])
        idx = np.nonzero(mask)[0]
    else:
        a = np.asarray(points)
        if a.dtype == bool:
            if a.shape != (body.n_points,):
                raise _arg_error("mask length must equal the number of points")
            idx = np.nonzero(a)[0]
        else:
            idx = a.astype(np.int64).reshape(-1)
            if idx.size and (idx.min() < 0 or idx.max() >= body.n_points):
                raise _arg_error("point index out of range")
    body.point_sets[name] = np.asarray(idx, dtype=np.int64)


def material(body, set_name=None, **kwargs):
    """material!(body; kwargs...) / material!(body, set_name; kwargs...): material_parameters.jl:64-92."""
    if set_name is None:
        body.point_params = []
        set_name = "all_points"
    _check_set(body, set_name)
    params = get_point_params(body.mat, kwargs)
    body.point_params.append(params)
    pts = body.point_sets[set_name]
    body.params_map[pts] = len(body.point_params)
    # set_failure_permissions!: fracture.jl:147-155
    body.fail_permit[pts] = params.has_fracture()


def failure_permit(body, set_name, permit):
    """failure_permit!: fracture.jl:38-42."""
    _check_set(body, set_name)
    body.fail_permit[body.point_sets[set_name]] = bool(permit)


def no_failure(body, set_name="all_points"):
    """no_failure!(body[, set_name]): fracture.jl:78-96."""
    _check_set(body, set_name)
    if np.any(body.params_map[body.point_sets[set_name]] == 0):
        raise _arg_error(f"Not all points of point set :{set_name} have material parameters!")
    failure_permit(body, set_name, False)


def precrack(body, set_a, set_b, update_dmg=True):
    """precrack!(body, set_a, set_b; update_dmg=true): predefined_cracks.jl:106-112."""
    _check_set(body, set_a)
    _check_set(body, set_b)
    if np.intersect1d(body.point_sets[set_a], body.point_sets[set_b]).size:
        raise _arg_error(f"the sets `{set_a}` and `{set_b}` intersect!")
    body.point_sets_precracks.append((set_a, set_b, not update_dmg))


def _conflict(body, field, set_name, dim):
    """has_boundary_condition_conflict: boundary_conditions.jl:288-321."""
    pts = body.point_sets[set_name]
    for lst in (body.single_dim_bcs, body.posdep_single_dim_bcs, body.pos_single_dim_bcs):
        for fld, s, d, _ in lst:
            if fld == field and d == dim and np.intersect1d(pts, body.point_sets[s]).size:
                return True
    return False


def _add_bc(body, lst, field, set_name, dim, f):
    if _conflict(body, field, set_name, dim):
        raise _arg_error("the specified condition conflicts with already existing conditions!")
    lst.append((field, set_name, dim, f))


def velocity_bc(f, body, set_name, dim):
    """velocity_bc!(f, body, set_name, dim): boundary_conditions.jl:422-435 (on velocity_half)."""
    _check_set(body, set_name)
    kind = condition_function_type(f)
    d = get_dim(dim)
    if kind == "sdbc":
        _add_bc(body, body.single_dim_bcs, "velocity_half", set_name, d, f)
    elif kind == "pdsdbc":
        _add_bc(body, body.posdep_single_dim_bcs, "velocity_half", set_name, d, f)
    else:
        raise _arg_error("velocity_bc! needs f(t) or f(p, t)")


def forcedensity_bc(f, body, set_name, dim):
    """forcedensity_bc!: boundary_conditions.jl:559-575 (on b_ext).  f(p) is the
    position-dependent form that the Newton-Krylov solver scales by the load factor."""
    _check_set(body, set_name)
    kind = condition_function_type(f)
    d = get_dim(dim)
    if kind == "sdbc":
        _add_bc(body, body.single_dim_bcs, "b_ext", set_name, d, f)
    elif kind == "pdsdbc":
        _add_bc(body, body.posdep_single_dim_bcs, "b_ext", set_name, d, f)
    else:
        _add_bc(body, body.pos_single_dim_bcs, "b_ext", set_name, d, f)


def displacement_bc(f, body, set_name, dim):
    """displacement_bc!(f, body, set_name, dim): boundary_conditions.jl:675-688."""
    _check_set(body, set_name)
    if condition_function_type(f) != "psdbc":
        raise _arg_error("the function for a displacement boundary condition must be of the form\n`f(p)`!")
    _add_bc(body, body.pos_single_dim_bcs, "displacement", set_name, get_dim(dim), f)


def _ic_conflict(body, field, set_name, dim):
    """has_initial_condition_conflict: initial_conditions.jl:123-131 (same field, same
    dimension, intersecting point sets)."""
    pts = body.point_sets[set_name]
    for lst in (body.single_dim_ics, body.posdep_single_dim_ics):
        for fld, s, d, _ in lst:
            if fld == field and d == dim and np.intersect1d(pts, body.point_sets[s]).size:
                return True
    return False


def velocity_ic(*args):
    """velocity_ic!(body, set_name, dim, value) / velocity_ic!(f, body, set_name, dim):
    initial_conditions.jl:178-193.  The velocity of the set's points before time
    integration; f(p) gets the reference position of a point, and a NaN value leaves the
    point alone (apply_ic!, :71-83).  Applied when the data handler is built, as
    apply_initial_conditions! is at chunk creation (threads_body_data_handler.jl:55)."""
    if len(args) != 4:
        raise _arg_error("velocity_ic(body, set_name, dim, value) or velocity_ic(f, body, set_name, dim)")
    if callable(args[0]):
        f, body, set_name, dim = args
        _check_set(body, set_name)
        try:
            names = [q.name for q in inspect.signature(f).parameters.values()]
        except (TypeError, ValueError):
            names = []
        if names != ["p"]:
            raise _arg_error("wrong arguments for position dependent initial condition function!\n"
                             "Initial conditions support only functions of type `f(p)`, where "
                             "`p` is the position vector [x, y, z] of each point in the point set!")
        lst, item = body.posdep_single_dim_ics, f
    else:
        body, set_name, dim, value = args
        _check_set(body, set_name)
        lst, item = body.single_dim_ics, float(value)
    d = get_dim(dim)
    if _ic_conflict(body, "velocity", set_name, d):
        raise _arg_error("the specified condition conflicts with already existing conditions!")
    lst.append(("velocity", set_name, d, item))


def initial_velocity(body):
    """The velocity field (n, 3) after apply_initial_conditions! (initial_conditions.jl:
    86-94): single-dim conditions first, then the position-dependent ones; zero elsewhere."""
    v = np.zeros((body.n_points, 3))
    for _, s, d, val in body.single_dim_ics:
        v[body.point_sets[s], d] = val
    for _, s, d, f in body.posdep_single_dim_ics:
        pts = body.point_sets[s]
        vals = eval_posdep(f, body.position[:, pts])
        keep = ~np.isnan(vals)
        v[pts[keep], d] = vals[keep]
    return v


def eval_posdep(f, X, t=None):
    """Evaluate f(p[, t]) for every column of X (3 x m): vectorised when the function
    accepts arrays, otherwise point by point as the reference does."""
    m = X.shape[1]
    args = (X,) if t is None else (X, t)
    try:
        v = np.asarray(f(*args), dtype=np.float64)
        v = np.broadcast_to(v, (m,)).astype(np.float64)
        # the vectorised call must agree with the per-point call on a few points
        for k in range(min(m, 3)):
            a = (X[:, k],) if t is None else (X[:, k], t)
            s = float(f(*a))
            if not (s == v[k] or (np.isnan(s) and np.isnan(v[k]))):
                raise ValueError
        return v
    except Exception:
        out = np.empty(m)
        for k in range(m):
            a = (X[:, k],) if t is None else (X[:, k], t)
            out[k] = float(f(*a))
        return out
