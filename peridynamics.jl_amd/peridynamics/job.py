"""Job and submit (src/core/job.jl:45-62, src/core/submit.jl:14-27, 99-115).

submit() runs the job on one GPU through MI355XBodyDataHandler; it is the sibling of the
reference's _submit_threads.  With a `path`, every `freq`-th step (and the reference
configuration, step 0) is written as VTK files under `<path>/vtk` through the asynchronous
export pipeline (peridynamics.vtk); without one, the exported fields are kept in memory
(job.results)."""
from .data_handler import MI355XBodyDataHandler
from .solvers import DynamicRelaxation, NewtonKrylov, VelocityVerlet
from .vtk import VTKExporter

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

    exporter = None
    if job.path and job.fields:
        exporter = VTKExporter(dh, job.path, job.fields, body_name=getattr(job.body, "name", ""))
        export = exporter.capture
    else:
        def export(dh_, n, t):
            if job.fields:
                job.results.append({"step": n, "time": t,
                                    **{f: dh_.get(f) for f in job.fields}})

    try:
        if isinstance(solver, VelocityVerlet):
            dh.init_time_solver(solver)
            dh.initialize(solver)
            if exporter is not None:
                exporter.capture(dh, 0, 0.0)  # export_reference_results
            dh.solve(solver, freq=job.freq if job.fields else 0, on_export=export)
        elif isinstance(solver, DynamicRelaxation):
            dh.calc_force_density(0.0, solver.dt)  # initialize!(dh, solver)
            if exporter is not None:
                exporter.capture(dh, 0, 0.0)
            dh.solve_dr(solver, freq=job.freq if job.fields else 0, on_export=export)
        elif isinstance(solver, NewtonKrylov):
            if exporter is not None:
                exporter.capture(dh, 0, 0.0)
            dh.solve_nk(solver, on_export=export if job.fields else None)
        else:
            raise TypeError(f"unsupported solver {type(solver).__name__}")
    finally:
        if exporter is not None:
            job.files = exporter.close()
    return dh
