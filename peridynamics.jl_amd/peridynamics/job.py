"""Job and submit (src/core/job.jl:45-62, src/core/submit.jl:14-27, 99-115).

submit() runs the job on one GPU through MI355XBodyDataHandler; it is the sibling of the
reference's _submit_threads.  Exported fields are kept in memory (job.results) instead of
being written as VTK files."""
from .data_handler import MI355XBodyDataHandler
from .solvers import DynamicRelaxation, NewtonKrylov, VelocityVerlet

DEFAULT_FIELDS = ("displacement", "damage")


class Job:
    def __init__(self, body, solver, path="", freq=10, fields=DEFAULT_FIELDS, device=0,
                 precision="fp64"):
        self.body = body
        self.solver = solver
        self.path = path
        self.freq = int(freq)
        self.fields = tuple(fields)
        self.device = device
        self.precision = precision
        self.results = []

    def __repr__(self):
        return f"Job({self.body!r}, {self.solver!r})"


def submit(job, quiet=True):
    """Run `job`; returns the data handler (its fields hold the final state)."""
    dh = MI355XBodyDataHandler(job.body, device=job.device, precision=job.precision)
    solver = job.solver

    def export(dh_, n, t):
        if job.fields:
            job.results.append({"step": n, "time": t,
                                **{f: dh_.get(f) for f in job.fields}})

    if isinstance(solver, VelocityVerlet):
        dh.init_time_solver(solver)
        dh.initialize(solver)
        dh.solve(solver, freq=job.freq if job.fields else 0, on_export=export)
    elif isinstance(solver, DynamicRelaxation):
        dh.calc_force_density(0.0, solver.dt)  # initialize!(dh, solver)
        dh.solve_dr(solver, freq=job.freq if job.fields else 0, on_export=export)
    elif isinstance(solver, NewtonKrylov):
        dh.solve_nk(solver, on_export=export if job.fields else None)
    else:
        raise TypeError(f"unsupported solver {type(solver).__name__}")
    return dh
