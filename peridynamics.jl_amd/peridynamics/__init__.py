"""MI355X-native force-density hot path of Peridynamics.jl, host side.

The reference's user API (Body, material!, velocity_bc!, VelocityVerlet, NewtonKrylov,
Job/submit) over libpdx.so, the C ABI of include/pdx.h.  Julia's `!` suffix is dropped
and point indices are 0-based."""
from ._lib import (ArgumentError, DuplicatePointError, NaNError, NotConvergedError, PdxError,
                   load as load_library)
from .body import (Body, displacement_bc, failure_permit, forcedensity_bc, material,
                   n_points, no_failure, point_set, precrack, velocity_bc, velocity_ic)
from .data_handler import MI355XBodyDataHandler
from .job import Job, submit
from .materials import (BBMaterial, CMaterial, CriticalStretch, EnergySurfaceCorrection,
                        LinearElastic, NeoHooke, NeoHookePenalty, NoCorrection, OSBMaterial,
                        SaintVenantKirchhoff, ZEMSilling, ZEMWan, const_one_kernel,
                        cubic_b_spline_kernel, cubic_b_spline_kernel_norm, linear_kernel)
from .point_generators import jitter, lattice_block, uniform_box
from .postprocess import crack_tip, mode_i_plate
from .solvers import DynamicRelaxation, NewtonKrylov, VelocityVerlet
from .vtk import VTKExporter, process_each_export, read_vtk, write_pvtu, write_vtu

__all__ = [
    "ArgumentError", "DuplicatePointError", "NaNError", "NotConvergedError", "PdxError",
    "load_library", "Body", "displacement_bc", "failure_permit", "forcedensity_bc", "material",
    "n_points", "no_failure", "point_set", "precrack", "velocity_bc", "velocity_ic",
    "MI355XBodyDataHandler",
    "Job", "submit", "BBMaterial", "CMaterial", "CriticalStretch", "NeoHookePenalty",
    "OSBMaterial", "SaintVenantKirchhoff", "LinearElastic", "NeoHooke", "ZEMSilling", "ZEMWan",
    "linear_kernel", "const_one_kernel", "cubic_b_spline_kernel", "cubic_b_spline_kernel_norm",
    "jitter",
    "lattice_block", "uniform_box", "NewtonKrylov", "VelocityVerlet", "crack_tip",
    "mode_i_plate", "DynamicRelaxation", "EnergySurfaceCorrection", "NoCorrection",
    "VTKExporter", "process_each_export", "read_vtk", "write_pvtu", "write_vtu",
]
