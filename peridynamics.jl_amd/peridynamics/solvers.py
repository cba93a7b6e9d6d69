"""Time solvers: VelocityVerlet (src/time_solvers/velocity_verlet.jl:47-163) and
NewtonKrylov (src/time_solvers/newton_krylov.jl:37-127, 222-295)."""
import math
import warnings

from ._lib import ArgumentError


def _arg_error(msg):
    return ArgumentError(-1, msg)


class VelocityVerlet:
    """VelocityVerlet(; time, steps, stepsize, safety_factor=0.7): velocity_verlet.jl:47-71."""

    def __init__(self, time=-1, steps=-1, stepsize=-1, safety_factor=0.7):
        if time > 0 and steps > 0:
            raise _arg_error("specify either time or number of steps, not both!")
        if time < 0 and steps < 0:
            raise _arg_error("specify either time or number of steps!")
        if not (0 < safety_factor < 1):
            raise _arg_error("wrong safety factor specified! condition: 0 < safety_factor < 1")
        if stepsize > 0:
            warnings.warn("stepsize specified! Please be sure that the CFD-condition holds!")
        self.end_time = float(time)
        self.n_steps = int(steps)
        self.dt = float(stepsize)
        self.safety_factor = float(safety_factor)

    def check(self):
        """velocity_verlet_check: velocity_verlet.jl:116-130."""
        if self.end_time < 0:
            raise RuntimeError("`end_time` of VelocityVerlet smaller than zero!\n")
        if self.n_steps < 0:
            raise RuntimeError("`n_steps` of VelocityVerlet smaller than zero!\n")
        if self.dt < 0:
            raise RuntimeError("`Δt` of VelocityVerlet smaller than zero!\n")
        if not (0 < self.safety_factor < 1):
            raise RuntimeError("`safety_factor` of VelocityVerlet has invalid value!\n")

    def __repr__(self):
        return f"VelocityVerlet(n_steps={self.n_steps}, safety_factor={self.safety_factor})"


class NewtonKrylov:
    """NewtonKrylov(; kwargs...): newton_krylov.jl:55-127.

    Additions of the MI355X build (defaults keep the reference behaviour):
    rel_tol > 0 stops a load step at ||r|| < rel_tol * ||r_first|| (the reference's `tol`
    is absolute, newton_krylov.jl:311-314); precond = "none" | "jacobi" | "two_level" | "learned"
    ("two_level": block Jacobi plus a Galerkin coarse correction by the translations of
    Morton-block aggregates; "learned": the same with a coarse space that also spans the
    trained message-passing network's fields).  With "learned", load steps 1 .. learn_steps
    are solved with the two-level preconditioner while the errors it leaves on their Krylov
    pairs are collected (at most krylov_pairs samples); the network is then trained
    (train_steps Adam steps on random subsets of train_batch samples, step size train_lr)
    and the remaining load steps use its coarse space.  The last tenth of the samples is held
    out; the fields are used only if their held-out loss is below `gate` times that of the
    translations alone (library default 0.95), otherwise the remaining steps run with the
    two-level preconditioner alone; so they do after the first learned solve that needs the
    fallback."""

    def __init__(self, time=-1, steps=-1, stepsize=1.0, maxiter=100, tol=1e-4,
                 perturbation_scale=1.0, gmres_maxiter=200, gmres_reltol=1e-4,
                 gmres_abstol=1e-8, gmres_restart=100, rel_tol=0.0, precond="none",
                 learn_steps=3, train_steps=40, train_batch=128, train_lr=6e-3,
                 krylov_pairs=512, train_seed=0, collect_precond="two_level", gate=None):
        if time > 0 and steps > 0:
            raise _arg_error("specify either time or number of steps, not both!")
        if time < 0 and steps < 0:
            raise _arg_error("specify either time or number of steps!")
        if stepsize <= 0:
            raise _arg_error("stepsize has to be larger than zero: stepsize > 0")
        if time > 0:
            n_steps = int(math.ceil(time / stepsize))
            end_time = float(time)
        else:
            n_steps = int(steps)
            end_time = steps * stepsize
        for name, val in (("maxiter", maxiter), ("tol", tol), ("gmres_reltol", gmres_reltol),
                          ("gmres_abstol", gmres_abstol), ("gmres_maxiter", gmres_maxiter),
                          ("gmres_restart", gmres_restart),
                          ("perturbation_scale", perturbation_scale)):
            if val <= 0:
                raise _arg_error(f"{name} has to be larger than zero: {name} > 0")
        if precond not in ("none", "jacobi", "two_level", "learned"):
            raise _arg_error("precond must be 'none', 'jacobi', 'two_level' or 'learned'")
        self.end_time = end_time
        self.n_steps = n_steps
        self.dt = float(stepsize)
        self.maxiter = int(maxiter)
        self.tol = float(tol)
        self.perturbation_scale = float(perturbation_scale)
        self.gmres_maxiter = int(gmres_maxiter)
        self.gmres_reltol = float(gmres_reltol)
        self.gmres_abstol = float(gmres_abstol)
        self.gmres_restart = int(gmres_restart)
        self.rel_tol = float(rel_tol)
        self.precond = precond
        self.learn_steps = int(learn_steps)
        self.train_steps = int(train_steps)
        self.train_batch = int(train_batch)
        self.train_lr = float(train_lr)
        self.krylov_pairs = int(krylov_pairs)
        self.train_seed = int(train_seed)
        if collect_precond != "two_level":
            raise _arg_error("collect_precond must be 'two_level' (the samples are its errors)")
        self.collect_precond = collect_precond
        self.gate = None if gate is None else float(gate)

    def __repr__(self):
        return f"NewtonKrylov(n_steps={self.n_steps}, stepsize={self.dt})"


class DynamicRelaxation:
    """DynamicRelaxation(; steps, stepsize=1.0, damping_factor=1.0):
    dynamic_relaxation.jl:36-58 (BB, OSB and CPD bodies on the device)."""

    def __init__(self, steps, stepsize=1.0, damping_factor=1.0):
        if steps <= 0:
            raise _arg_error("`steps` should be larger than zero!\n")
        if stepsize <= 0:
            raise _arg_error("`stepsize` should be larger than zero!\n")
        if damping_factor <= 0:
            raise _arg_error("`damping_factor` should be larger than zero!\n")
        self.n_steps = int(steps)
        self.dt = float(stepsize)
        self.damping = float(damping_factor)

    def __repr__(self):
        return f"DynamicRelaxation(n_steps={self.n_steps}, Δt={self.dt}, Λ={self.damping})"
