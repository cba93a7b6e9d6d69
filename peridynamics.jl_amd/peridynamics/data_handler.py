"""MI355XBodyDataHandler: one body, resident in the HBM of one GPU, behind libpdx.

It takes the place of the reference's ThreadsBodyDataHandler
(src/core/threads_body_data_handler.jl:24-39, 184-212) for a single body: there is one
chunk, no halo exchange, and every per-step sweep runs inside libpdx.  The host keeps
what the reference keeps on the host side of a solver loop: the boundary-condition
closures, which it evaluates per step and hands to the device as plain values."""
import numpy as np

from . import _lib
from ._lib import ArgumentError, Context, NaNError, PdxNkOpts, PdxNkStats, PdxParams
from .body import eval_posdep, initial_velocity
from .materials import KERNEL_IDS, CMaterial, OSBMaterial


def _arg_error(msg):
    return ArgumentError(-1, msg)


PRECISIONS = {"fp64": _lib.PDX_FP64, "mixed": _lib.PDX_MIXED}


class MI355XBodyDataHandler:
    def __init__(self, body, device=0, precision="fp64"):
        if precision not in PRECISIONS:
            raise _arg_error("precision must be 'fp64' or 'mixed'")
        if np.any(body.params_map == 0):
            raise _arg_error("not all points have material parameters (material!)")
        # distinct parameter sets in use (ParameterHandler, parameter_handler.jl:36-55)
        used = sorted(set(int(k) for k in body.params_map))
        keys = []
        set_of_key = {}
        for k in used:
            key = body.point_params[k - 1].key()
            if key not in set_of_key:
                set_of_key[key] = len(keys)
                keys.append(k)
        self.body = body
        self.params = body.point_params[keys[0] - 1]
        self.precision = precision
        self.ctx = Context(device, PRECISIONS[precision])
        self.n = body.n_points
        self.ctx.set_points(body.xyz, body.volume, body.fail_permit.astype(np.uint8))
        mat = body.mat
        sets = [self._pdx_params(body.point_params[k - 1], mat) for k in keys]
        self.pp = sets[0]
        if len(sets) == 1:
            self.ctx.set_material(mat.kind, self.pp)
        else:
            remap = np.zeros(len(body.point_params) + 1, np.int32)
            for k in used:
                remap[k] = set_of_key[body.point_params[k - 1].key()]
            self.ctx.set_material_sets(mat.kind, sets, remap[body.params_map])
        self.n_sets = len(sets)
        corr = getattr(mat, "correction", None)
        if corr is not None and corr.code:
            self.ctx.set_surface_correction(corr.code)
        for a, b, filt in body.point_sets_precracks:
            self.ctx.add_precrack(body.point_sets[a], body.point_sets[b], filt)
        self.n_bonds = self.ctx.build_bonds()
        if body.single_dim_ics or body.posdep_single_dim_ics:
            self.ctx.set_field("velocity", initial_velocity(body))
        self._compile_bcs()
        self.dt = None

    @staticmethod
    def _pdx_params(p, mat):
        pp = PdxParams(delta=p.delta, rho=p.rho, E=p.E, nu=p.nu, G=p.G, K=p.K, lambda_=p.lam,
                       mu=p.mu, Gc=p.Gc, epsc=p.epsc, bc=p.bc, model=0, kernel=0, zem_cs=0.5,
                       maxdmg=0.85)
        if isinstance(mat, (OSBMaterial, CMaterial)):
            pp.kernel = KERNEL_IDS[mat.kernel]
        if isinstance(mat, CMaterial):
            pp.model = mat.constitutive_model.code
            pp.zem = mat.zem.code
            pp.zem_cs = getattr(mat.zem, "Cs", 0.0)
            pp.maxdmg = mat.maxdmg
        return pp

    # ------------------------------------------------------------------ boundary conditions
    def _compile_bcs(self):
        """velocity_bc! and time-dependent forcedensity_bc! conditions: one source per
        single-dim condition and per point of a position-dependent one, evaluated by the
        host for every step (bc_values) and applied on the device."""
        body = self.body
        conds = {"velocity_half": ([], []), "b_ext": ([], [])}
        for fld, s, d, f in body.single_dim_bcs:
            if fld in conds:
                conds[fld][0].append((s, d, f))
        for fld, s, d, f in body.posdep_single_dim_bcs:
            if fld in conds:
                conds[fld][1].append((s, d, f))
        n_src = 0
        self._sd = []
        self._pd = []
        tables = {}
        for fld, (sd, pd) in conds.items():
            sets = [body.point_sets[s] for s, _, _ in sd + pd]
            pts = np.unique(np.concatenate(sets)) if sets else np.zeros(0, np.int64)
            slot = {int(p): k for k, p in enumerate(pts)}
            src = -np.ones((len(pts), 3), np.int32)
            for s, d, f in sd:
                idx = [slot[int(p)] for p in body.point_sets[s]]
                src[idx, d] = n_src
                self._sd.append((n_src, f))
                n_src += 1
            for s, d, f in pd:
                ps = body.point_sets[s]
                idx = [slot[int(p)] for p in ps]
                src[idx, d] = n_src + np.arange(len(ps))
                self._pd.append((n_src, ps, f))
                n_src += len(ps)
            tables[fld] = (pts, src)
        self.n_src = n_src
        pts, src = tables["velocity_half"]
        self.ctx.set_velocity_bcs(pts.astype(np.int32), src, n_src)
        pts, src = tables["b_ext"]
        if len(pts):
            self.ctx.set_force_density_bcs(pts.astype(np.int32), src)

    def bc_values(self, steps, dt):
        """Values of every condition source for each step n (time n*dt)."""
        steps = list(steps)
        vals = np.empty((len(steps), max(self.n_src, 0)), np.float64)
        X = self.body.position
        for k, n in enumerate(steps):
            t = n * dt
            for s, f in self._sd:
                vals[k, s] = float(f(t))
            for base, ps, f in self._pd:
                vals[k, base:base + len(ps)] = eval_posdep(f, X[:, ps], t)
        return vals

    # ------------------------------------------------------------------ reference interface
    def calc_force_density(self, t=0.0, dt=0.0):
        """calc_force_density!(dh, t, Δt): threads_body_data_handler.jl:184-195."""
        self.ctx.force_density(t, dt)

    def calc_stable_timestep(self, safety_factor=0.7):
        """calc_stable_timestep(dh, safety_factor): velocity_verlet.jl:136-163."""
        return self.ctx.stable_timestep(safety_factor)

    def init_time_solver(self, vv):
        """init_time_solver!(vv, dh): velocity_verlet.jl:103-114."""
        if vv.dt < 0:
            vv.dt = self.calc_stable_timestep(vv.safety_factor)
        if vv.end_time < 0:
            vv.end_time = vv.n_steps * vv.dt
        elif vv.n_steps < 0:
            vv.n_steps = int(vv.end_time // vv.dt + 1)
        vv.check()
        self.dt = vv.dt

    def initialize(self, vv):
        """initialize!(dh, solver): data_handler.jl:1-4 (force at t = 0; a stays 0)."""
        self.calc_force_density(0.0, vv.dt)

    def solve(self, vv, freq=0, on_export=None, first_step=1, max_chunk=None):
        """solve!(dh, vv, options): velocity_verlet.jl:165-179.  on_export(dh, n, t) runs
        every `freq` steps (export_results)."""
        dt = vv.dt
        remaining = vv.n_steps - first_step + 1
        n = first_step
        cap = max(1, (1 << 24) // max(self.n_src, 1))
        if max_chunk:
            cap = min(cap, max_chunk)
        if freq and freq > 0:
            cap = min(cap, freq)
        while remaining > 0:
            k = min(cap, remaining)
            if freq and freq > 0:
                k = min(k, freq - ((n - 1) % freq))
            vals = self.bc_values(range(n, n + k), dt) if self.n_src else None
            try:
                self.ctx.vv_run(n, k, dt, vals)
            except NaNError as e:
                raise NaNError(e.status, f"NaN in the force density at t={e.step * dt}, step {e.step}",
                               step=e.step, t=e.step * dt) from None
            n += k
            remaining -= k
            if on_export is not None and freq and (n - 1) % freq == 0:
                on_export(self, n - 1, (n - 1) * dt)
        return self

    # ------------------------------------------------------------------ dynamic relaxation
    def solve_dr(self, dr, freq=0, on_export=None):
        """solve!(dh, dr, options): dynamic_relaxation.jl:92-107 (init_density_matrix!, then
        relaxation_timestep! for n = 1 .. n_steps).  on_export(dh, n, t) every `freq` steps."""
        dt = dr.dt
        self.ctx.dr_init(dt, dr.damping)
        n, remaining = 1, dr.n_steps
        cap = max(1, (1 << 24) // max(self.n_src, 1))
        if freq and freq > 0:
            cap = min(cap, freq)
        while remaining > 0:
            k = min(cap, remaining)
            if freq and freq > 0:
                k = min(k, freq - ((n - 1) % freq))
            vals = self.bc_values(range(n, n + k), dt) if self.n_src else None
            try:
                self.ctx.dr_run(n, k, dt, vals)
            except NaNError as e:
                raise NaNError(e.status, f"NaN in the force density at t={e.step * dt}, step {e.step}",
                               step=e.step, t=e.step * dt) from None
            n += k
            remaining -= k
            if on_export is not None and freq and (n - 1) % freq == 0:
                on_export(self, n - 1, (n - 1) * dt)
        return self

    def calc_damping(self, dt):
        """calc_damping(dh, dt) of the current fields: dynamic_relaxation.jl:262-323."""
        return self.ctx.dr_damping(dt)

    # ------------------------------------------------------------------ Newton-Krylov
    PRECONDS = {"none": 0, "jacobi": 1, "learned": 2, "two_level": 3}

    def nk_options(self, nk, precond=None):
        pre = self.PRECONDS[precond or nk.precond]
        return PdxNkOpts(maxiter=nk.maxiter, gmres_maxiter=nk.gmres_maxiter,
                         gmres_restart=nk.gmres_restart, precond=pre,
                         tol=nk.tol, rel_tol=nk.rel_tol, gmres_reltol=nk.gmres_reltol,
                         gmres_abstol=nk.gmres_abstol,
                         perturbation_scale=nk.perturbation_scale)

    def init_newton_krylov(self, nk):
        """validate_chunk_for_newton_krylov! (newton_krylov.jl:222-256) and the constrained
        dofs of displacement_bc! (full-load values f(X), scaled per load step)."""
        body = self.body
        if body.single_dim_bcs or body.posdep_single_dim_bcs:
            raise _arg_error("NewtonKrylov solver does not support all boundary condition types!\n"
                             "Specify only position-dependent boundary conditions!")
        for k in sorted(set(int(k) for k in body.params_map)):
            if body.point_params[k - 1].has_fracture():
                i = int(np.argmax(body.params_map == k))
                raise _arg_error("NewtonKrylov solver currently does not support damage or "
                                 f"fracture!\nFracture parameters found at point index {i + 1}!")
        pts, dims, vals = [], [], []
        X = body.position
        for fld, s, d, f in body.pos_single_dim_bcs:
            if fld != "displacement":
                continue
            P = body.point_sets[s]
            pts.append(P)
            dims.append(np.full(len(P), d, np.int32))
            vals.append(eval_posdep(f, X[:, P]))
        if pts:
            self.ctx.set_displacement_bcs(np.concatenate(pts), np.concatenate(dims),
                                          np.concatenate(vals))
        else:
            self.ctx.set_displacement_bcs(np.zeros(0, np.int32), np.zeros(0, np.int32),
                                          np.zeros(0))
        self._bext_pos = [(body.point_sets[s], d, f) for fld, s, d, f in body.pos_single_dim_bcs
                          if fld == "b_ext"]

    def nk_load_step(self, nk, n, precond=None):
        """newton_krylov_step!(dh, nk, options, n): newton_krylov.jl:298-332."""
        beta = n / nk.n_steps
        if self._bext_pos:
            bext = np.zeros((self.n, 3))
            X = self.body.position
            for P, d, f in self._bext_pos:
                val = beta * eval_posdep(f, X[:, P])
                ok = np.isfinite(val)
                bext[P[ok], d] = val[ok]
            self.set("b_ext", bext)
        return self.ctx.nk_step(n, nk.n_steps, self.nk_options(nk, precond))

    def solve_nk(self, nk, on_export=None):
        """solve!(dh, nk, options): newton_krylov.jl:280-295.  Returns the per-step stats."""
        self.init_newton_krylov(nk)
        stats = []
        self.train_info = None
        learned = nk.precond == "learned"
        for n in range(1, nk.n_steps + 1):
            if learned and n <= nk.learn_steps:
                # plain (Jacobi) solves that supply the training pairs
                self.ctx.gnn_collect(True, nk.krylov_pairs)
                stats.append(self.nk_load_step(nk, n, precond=nk.collect_precond))
                self.ctx.gnn_collect(False)
                if n == min(nk.learn_steps, nk.n_steps):
                    self.train_learned(nk)
            else:
                if learned and nk.learn_steps <= 0 and self.train_info is None:
                    raise _arg_error("the learned preconditioner needs learn_steps >= 1")
                stats.append(self.nk_load_step(nk, n))
            if on_export is not None:
                on_export(self, n, n * nk.dt)
        return stats

    def train_learned(self, nk):
        """Trains the message-passing preconditioner on the collected pairs."""
        import time
        t0 = time.perf_counter()
        pairs = self.ctx.gnn_pairs()
        loss = self.ctx.gnn_train(nk.train_steps, nk.train_batch, nk.train_lr, nk.train_seed)
        val_loss, use_net = self.ctx.gnn_gate(nk.gate if nk.gate is not None else float("nan"))
        self.ctx.synchronize()
        self.train_info = {"pairs": pairs, "loss": loss, "val_loss": val_loss,
                           "network_applied": use_net, "seconds": time.perf_counter() - t0}
        return self.train_info

    # ------------------------------------------------------------------ fields
    def get(self, name):
        """A storage field in the caller's point order, shape (n, 3), (n, 9) or (n,)."""
        return self.ctx.get_field(name)

    def set(self, name, values):
        self.ctx.set_field(name, values)

    def bonds(self):
        """(row_ptr, nbr, L, fail_permit, active) in the caller's point order, 0-based."""
        return self.ctx.get_bonds()

    def surface_correction(self):
        """EnergySurfaceCorrection factors: (scfactor per bond in the order of bonds(),
        mfactor (n, 3)); system.correction.scfactor / .mfactor of the reference."""
        return self.ctx.get_surface_correction(self.n_bonds)

    def close(self):
        self.ctx.close()
