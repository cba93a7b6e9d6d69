# PeridynamicsMI355X.jl -- Julia side of the MI355X force-density library (libpdx.so).
#
# A data handler for Peridynamics.jl that keeps the whole body on one GPU and routes the
# hot path through the C ABI declared in include/pdx.h.  The package's own API stays as it
# is (Body, material!, velocity_bc!, displacement_bc!, precrack!, VelocityVerlet,
# NewtonKrylov, Job); `submit_mi355x(job)` is the sibling of `_submit_threads`
# (src/core/submit.jl:99-115) that builds an `MI355XBodyDataHandler` instead of
# `threads_data_handler` (src/core/threads_body_data_handler.jl:31-60).
#
# Methods added to the package's generic functions:
#   calc_force_density!(dh, t, Δt)      src/core/threads_body_data_handler.jl:184-195
#   calc_stable_timestep(dh, sf)        src/time_solvers/velocity_verlet.jl:136-163
#   initialize!(dh, solver)             src/core/data_handler.jl:1-4
#   solve!(dh, ::VelocityVerlet, opts)  src/time_solvers/velocity_verlet.jl:165-197
#   solve!(dh, ::DynamicRelaxation, o)  src/time_solvers/dynamic_relaxation.jl:92-107
#   solve!(dh, ::NewtonKrylov, opts)    src/time_solvers/newton_krylov.jl:280-332
#   export_results / export_reference_results  src/core/threads_body_data_handler.jl:197-212
#     (device snapshot through pdx_export_begin/wait, files by the package's own
#      pvtk_grid writer as _export_results does, src/auxiliary/io.jl:157-167)
#
# Status codes map to the package's exceptions: PDX_ERR_ARG -> ArgumentError,
# PDX_ERR_NAN -> NaNError(t, step) (src/auxiliary/errors.jl:98-114), everything else ->
# ErrorException with pdx_last_error().
module PeridynamicsMI355X

using Peridynamics
import Peridynamics: calc_force_density!, calc_stable_timestep, initialize!, solve!,
                     init_time_solver!, AbstractDataHandler, AbstractJobOptions, NaNError,
                     EnergySurfaceCorrection, correction_type, get_filename
using WriteVTK: pvtk_grid, MeshCell, VTKCellTypes, VTKPointData, VTKFieldData

export MI355XBodyDataHandler, submit_mi355x, get_field, set_field!

const libpdx = get(ENV, "PDX_LIB", joinpath(@__DIR__, "..", "libpdx.so"))

const PDX_OK = Int32(0)
const PDX_ERR_ARG = Int32(-1)
const PDX_ERR_NAN = Int32(-4)
const PDX_ERR_NOT_CONVERGED = Int32(-7)
const PDX_FP64 = Int32(0)
const PDX_MIXED = Int32(1)

const FIELDS = Dict(:position => (0, 3), :displacement => (1, 3), :velocity => (2, 3),
                    :velocity_half => (3, 3), :acceleration => (4, 3), :b_int => (5, 3),
                    :b_ext => (6, 3), :damage => (7, 1), :n_active_bonds => (8, 1),
                    :defgrad => (9, 9), :cauchy_stress => (10, 9))

# struct layouts of include/pdx.h
struct PdxParams
    delta::Float64; rho::Float64; E::Float64; nu::Float64; G::Float64; K::Float64
    lambda::Float64; mu::Float64; Gc::Float64; epsc::Float64; bc::Float64
    model::Int32; kernel::Int32
    zem_cs::Float64; maxdmg::Float64
    zem::Int32; reserved_::Int32
end

struct PdxNkOpts
    maxiter::Int32; gmres_maxiter::Int32; gmres_restart::Int32; precond::Int32
    tol::Float64; rel_tol::Float64; gmres_reltol::Float64; gmres_abstol::Float64
    perturbation_scale::Float64
end

mutable struct PdxNkStats
    newton_iters::Int32; krylov_iters::Int32; jv_evals::Int32; precond_fallbacks::Int32
    residual_first::Float64; residual_final::Float64; seconds::Float64
    PdxNkStats() = new(0, 0, 0, 0, 0.0, 0.0, 0.0)
end

mutable struct MI355XBodyDataHandler{B<:AbstractBody} <: AbstractDataHandler
    body::B
    ctx::Ptr{Cvoid}
    n_points::Int
    n_bonds::Int
    # velocity_bc! and time-dependent forcedensity_bc! tables: the host evaluates the
    # closures once per step, the value row holds one source per SingleDimBC and one per
    # point of a PosDepSingleDimBC
    sd::Vector{Any}        # (source, SingleDimBC) on :velocity_half and :b_ext
    pd::Vector{Any}        # (first source, PosDepSingleDimBC) on :velocity_half and :b_ext
    pd_points::Vector{Vector{Int}}
    n_src::Int
end

last_error(ctx) = unsafe_string(ccall((:pdx_last_error, libpdx), Cstring, (Ptr{Cvoid},), ctx))

function check(dh_or_ctx, st::Int32; t=NaN, step=0)
    st == PDX_OK && return nothing
    ctx = dh_or_ctx isa MI355XBodyDataHandler ? dh_or_ctx.ctx : dh_or_ctx
    msg = last_error(ctx)
    st == PDX_ERR_ARG && throw(ArgumentError(msg))
    st == PDX_ERR_NAN && throw(NaNError(t, step))
    error(msg)
end

material_kind(::BBMaterial) = Int32(0)
material_kind(::OSBMaterial) = Int32(1)
material_kind(::CMaterial) = Int32(2)
material_kind(m) = throw(ArgumentError("$(typeof(m)) has no MI355X kernel"))

model_id(::CMaterial{<:SaintVenantKirchhoff}) = Int32(0)
model_id(::CMaterial{<:NeoHookePenalty}) = Int32(1)
model_id(::CMaterial{<:LinearElastic}) = Int32(2)
model_id(::CMaterial{<:NeoHooke}) = Int32(3)
model_id(_) = Int32(0)

# influence function of OSBMaterial / CMaterial: the four of src/discretization/kernels.jl
# (PDX_KERNEL_* of pdx.h); a user function has no device implementation and is refused, not
# replaced.  BBMaterial has none.
kernel_id(::BBMaterial) = Int32(0)
function kernel_id(mat::Union{OSBMaterial,CMaterial})
    mat.kernel === linear_kernel && return Int32(0)
    mat.kernel === const_one_kernel && return Int32(1)
    mat.kernel === cubic_b_spline_kernel && return Int32(2)
    mat.kernel === cubic_b_spline_kernel_norm && return Int32(3)
    throw(ArgumentError("kernel $(mat.kernel) has no MI355X implementation; the device " *
                        "evaluates linear_kernel, const_one_kernel, cubic_b_spline_kernel " *
                        "and cubic_b_spline_kernel_norm"))
end

function pdx_params(body, p)
    wan = body.mat isa CMaterial && body.mat.zem isa ZEMWan
    cs = body.mat isa CMaterial && !wan ? body.mat.zem.Cs : 0.0
    maxdmg = body.mat isa CMaterial ? body.mat.maxdmg : 0.0
    return PdxParams(p.δ, p.rho, p.E, p.nu, p.G, p.K, p.λ, p.μ, p.Gc, p.εc, p.bc,
                     model_id(body.mat), kernel_id(body.mat), cs, maxdmg, Int32(wan ? 1 : 0),
                     Int32(0))
end

# one parameter set: pdx_set_material; several (point_params + params_map):
# pdx_set_material_sets with the 0-based set of every point
function set_material!(ctx, body)
    sets = [pdx_params(body, p) for p in body.point_params]
    if length(sets) == 1
        check(ctx, ccall((:pdx_set_material, libpdx), Int32, (Ptr{Cvoid}, Int32, Ref{PdxParams}),
                         ctx, material_kind(body.mat), Ref(sets[1])))
    else
        check(ctx, ccall((:pdx_set_material_sets, libpdx), Int32,
                         (Ptr{Cvoid}, Int32, Int32, Ptr{PdxParams}, Ptr{Int32}),
                         ctx, material_kind(body.mat), length(sets), sets,
                         Int32.(body.params_map .- 1)))
    end
    if body.mat isa Union{BBMaterial,OSBMaterial} &&
       correction_type(body.mat) === EnergySurfaceCorrection
        check(ctx, ccall((:pdx_set_surface_correction, libpdx), Int32, (Ptr{Cvoid}, Int32),
                         ctx, Int32(1)))
    end
    return nothing
end

zero_based(v) = Int32.(v .- 1)

function MI355XBodyDataHandler(body::AbstractBody; device::Int=0, precision::Symbol=:fp64)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    prec = precision === :mixed ? PDX_MIXED : PDX_FP64
    st = ccall((:pdx_create, libpdx), Int32, (Ptr{Ptr{Cvoid}}, Int32, Int32), r, device, prec)
    st == PDX_OK || error("pdx_create failed ($st)")
    ctx = r[]
    n = body.n_points
    fp = UInt8.(body.fail_permit)
    check(ctx, ccall((:pdx_set_points, libpdx), Int32,
                     (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{UInt8}),
                     ctx, n, body.position, body.volume, fp))
    set_material!(ctx, body)
    for pc in body.point_sets_precracks
        a = zero_based(body.point_sets[pc.set_a])
        b = zero_based(body.point_sets[pc.set_b])
        check(ctx, ccall((:pdx_add_precrack, libpdx), Int32,
                         (Ptr{Cvoid}, Int64, Ptr{Int32}, Int64, Ptr{Int32}, Int32),
                         ctx, length(a), a, length(b), b, Int32(pc.filter_bonds)))
    end
    nb = Ref{Int64}(0)
    check(ctx, ccall((:pdx_build_bonds, libpdx), Int32, (Ptr{Cvoid}, Ref{Int64}), ctx, nb))
    dh = MI355XBodyDataHandler(body, ctx, n, Int(nb[]), Any[], Any[], Vector{Int}[], 0)
    finalizer(dh) do d
        d.ctx == C_NULL || ccall((:pdx_destroy, libpdx), Cvoid, (Ptr{Cvoid},), d.ctx)
        d.ctx = C_NULL
    end
    apply_initial_conditions!(dh)
    set_step_bcs!(dh)
    return dh
end

# apply_initial_conditions! (src/physics/initial_conditions.jl:86-94, called at chunk
# creation, threads_body_data_handler.jl:55): velocity_ic! values, then the f(p) ones
# (a NaN leaves the point alone), written to the device's velocity field
function apply_initial_conditions!(dh::MI355XBodyDataHandler)
    body = dh.body
    isempty(body.single_dim_ics) && isempty(body.posdep_single_dim_ics) && return nothing
    fields = Dict{Symbol,Matrix{Float64}}()
    field(f) = get!(() -> zeros(3, dh.n_points), fields, f)
    for ic in body.single_dim_ics
        v = field(ic.field)
        for i in body.point_sets[ic.point_set]
            v[ic.dim, i] = ic.value
        end
    end
    for ic in body.posdep_single_dim_ics
        v = field(ic.field)
        for i in body.point_sets[ic.point_set]
            value = ic.fun(view(body.position, :, i))
            isnan(value) || (v[ic.dim, i] = value)
        end
    end
    for (f, v) in fields
        haskey(FIELDS, f) ||
            throw(ArgumentError("initial condition on `$f` has no MI355X field"))
        set_field!(dh, f, v)
    end
    return nothing
end

function get_field(dh::MI355XBodyDataHandler, name::Symbol)
    id, nc = FIELDS[name]
    out = nc == 1 ? zeros(dh.n_points) : zeros(nc, dh.n_points)
    check(dh, ccall((:pdx_get_field, libpdx), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}),
                    dh.ctx, id, out))
    return out
end

function set_field!(dh::MI355XBodyDataHandler, name::Symbol, values::AbstractArray)
    id, _ = FIELDS[name]
    check(dh, ccall((:pdx_set_field, libpdx), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}),
                    dh.ctx, id, Float64.(values)))
    return nothing
end

# velocity_bc! (boundary_conditions.jl:422-435) on velocity_half and forcedensity_bc! with
# f(t) or f(p, t) (:559-575) on b_ext: single-dim BCs first, then the position-dependent
# ones, as apply_boundary_conditions! orders them (:212-226).  One value row per step holds
# the sources of both fields (pdx_set_velocity_bcs takes the row length, then
# pdx_set_force_density_bcs indexes the same rows).  Conditions on any other field, and
# data-driven conditions, are refused rather than dropped.
function set_step_bcs!(dh::MI355XBodyDataHandler)
    body = dh.body
    isempty(body.data_bcs) ||
        throw(ArgumentError("data boundary conditions have no MI355X implementation"))
    for bc in vcat(body.single_dim_bcs, body.posdep_single_dim_bcs)
        bc.field in (:velocity_half, :b_ext) ||
            throw(ArgumentError("boundary condition on `$(bc.field)` has no MI355X implementation"))
    end
    sd_all, pd_all, pd_points = Any[], Any[], Vector{Int}[]
    tables = Dict{Symbol,Tuple{Vector{Int},Vector{Int32}}}()
    k = 0
    for fld in (:velocity_half, :b_ext)
        sd = [bc for bc in body.single_dim_bcs if bc.field === fld]
        pd = [bc for bc in body.posdep_single_dim_bcs if bc.field === fld]
        src = Dict{Int,NTuple{3,Int}}()
        put!(i, d, s) = (v = collect(get(src, i, (-1, -1, -1))); v[d] = s; src[i] = Tuple(v))
        for bc in sd
            for i in body.point_sets[bc.point_set]
                put!(i, Int(bc.dim), k)
            end
            push!(sd_all, (k, bc))
            k += 1
        end
        for bc in pd
            pts = body.point_sets[bc.point_set]
            push!(pd_all, (k, bc))
            push!(pd_points, pts)
            for (q, i) in enumerate(pts)
                put!(i, Int(bc.dim), k + q - 1)
            end
            k += length(pts)
        end
        points = sort!(collect(keys(src)))
        tables[fld] = (points, Int32[s for i in points for s in src[i]])
    end
    points, srcv = tables[:velocity_half]
    check(dh, ccall((:pdx_set_velocity_bcs, libpdx), Int32,
                    (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Int32}, Int64),
                    dh.ctx, length(points), zero_based(points), srcv, k))
    points, srcv = tables[:b_ext]
    if !isempty(points)
        check(dh, ccall((:pdx_set_force_density_bcs, libpdx), Int32,
                        (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Int32}),
                        dh.ctx, length(points), zero_based(points), srcv))
    end
    dh.sd, dh.pd, dh.pd_points, dh.n_src = sd_all, pd_all, pd_points, k
    return nothing
end

function bc_values(dh::MI355XBodyDataHandler, steps, Δt)
    vals = zeros(dh.n_src, length(steps))
    X = dh.body.position
    for (c, n) in enumerate(steps)
        t = n * Δt
        for (k, bc) in dh.sd
            vals[k + 1, c] = bc.fun(t)
        end
        for ((k0, bc), pts) in zip(dh.pd, dh.pd_points)
            for (q, i) in enumerate(pts)
                vals[k0 + q, c] = bc.fun(view(X, :, i), t)
            end
        end
    end
    return vals
end

function calc_force_density!(dh::MI355XBodyDataHandler, t, Δt)
    check(dh, ccall((:pdx_force_density, libpdx), Int32, (Ptr{Cvoid}, Float64, Float64),
                    dh.ctx, t, Δt); t=t)
    return nothing
end

function calc_stable_timestep(dh::MI355XBodyDataHandler, safety_factor::Float64)
    dt = Ref{Float64}(0.0)
    check(dh, ccall((:pdx_stable_timestep, libpdx), Int32, (Ptr{Cvoid}, Float64, Ref{Float64}),
                    dh.ctx, safety_factor, dt))
    return dt[]
end

initialize!(dh::MI355XBodyDataHandler, ::AbstractTimeSolver) = calc_force_density!(dh, 0.0, 0.0)

# export_results: the snapshot of step n is taken on the device and copied while the next
# steps run; the files of the previous snapshot are written meanwhile.  `pending` holds the
# (step, time) of the snapshot in flight.
const EXPORT_PENDING = IdDict{Any,Any}()

function export_fields(dh, options, n, t)
    options isa Peridynamics.JobOptions && options.export_allowed || return nothing
    names = vcat(:position, [f for f in options.fields if f !== :position])
    ids = Int32[FIELDS[f][1] for f in names]
    check(dh, ccall((:pdx_export_begin, libpdx), Int32, (Ptr{Cvoid}, Int32, Ptr{Int32}, Int64),
                    dh.ctx, length(ids), ids, n))
    prev = get(EXPORT_PENDING, dh, nothing)
    EXPORT_PENDING[dh] = (names, t)
    prev === nothing || write_export(dh, options, prev...)
    return nothing
end

function write_export(dh, options, names, t)
    ptrs = Vector{Ptr{Float64}}(undef, 16)
    tag, slot = Ref{Int64}(0), Ref{Int32}(0)
    check(dh, ccall((:pdx_export_wait, libpdx), Int32,
                    (Ptr{Cvoid}, Ref{Int64}, Ref{Int32}, Ptr{Ptr{Float64}}),
                    dh.ctx, tag, slot, ptrs))
    n = dh.n_points
    arr(k) = (nc = FIELDS[names[k]][2];
              nc == 1 ? copy(unsafe_wrap(Array, ptrs[k], n)) :
                        copy(unsafe_wrap(Array, ptrs[k], (nc, n))))
    cells = [MeshCell(VTKCellTypes.VTK_VERTEX, (j,)) for j in 1:n]
    filename = get_filename(options, Symbol(""), Int(tag[]))
    pvtk_grid(filename, arr(1), cells; part=1, nparts=1) do vtk
        for k in 2:length(names)
            vtk[string(names[k]), VTKPointData()] = arr(k)
        end
        vtk["time", VTKFieldData()] = t
    end
    return nothing
end

function finish_exports(dh, options)
    prev = pop!(EXPORT_PENDING, dh, nothing)
    prev === nothing || write_export(dh, options, prev...)
    return nothing
end

function solve!(dh::MI355XBodyDataHandler, vv::VelocityVerlet, options::AbstractJobOptions)
    Δt = vv.Δt
    freq = options isa Peridynamics.JobOptions && options.export_allowed ? options.freq : 0
    freq > 0 && export_fields(dh, options, 0, 0.0)  # export_reference_results
    n = 1
    while n <= vv.n_steps
        # run to the next export step (or the end) in one call: one fused launch per step
        last = freq > 0 ? min(vv.n_steps, (div(n - 1, freq) + 1) * freq) : vv.n_steps
        k = last - n + 1
        vals = bc_values(dh, n:last, Δt)
        nan_step = Ref{Int64}(0)
        st = ccall((:pdx_vv_run, libpdx), Int32,
                   (Ptr{Cvoid}, Int64, Int64, Float64, Ptr{Float64}, Ref{Int64}),
                   dh.ctx, n, k, Δt, vals, nan_step)
        check(dh, st; t=nan_step[] * Δt, step=nan_step[])
        freq > 0 && last % freq == 0 && export_fields(dh, options, last, last * Δt)
        n = last + 1
    end
    finish_exports(dh, options)
    return dh
end

function solve!(dh::MI355XBodyDataHandler, dr::DynamicRelaxation, options::AbstractJobOptions)
    Δt = dr.Δt
    check(dh, ccall((:pdx_dr_init, libpdx), Int32, (Ptr{Cvoid}, Float64, Float64),
                    dh.ctx, Δt, dr.Λ))
    freq = options isa Peridynamics.JobOptions && options.export_allowed ? options.freq : 0
    freq > 0 && export_fields(dh, options, 0, 0.0)
    n = 1
    while n <= dr.n_steps
        last = freq > 0 ? min(dr.n_steps, (div(n - 1, freq) + 1) * freq) : dr.n_steps
        vals = bc_values(dh, n:last, Δt)
        nan_step = Ref{Int64}(0)
        st = ccall((:pdx_dr_run, libpdx), Int32,
                   (Ptr{Cvoid}, Int64, Int64, Float64, Ptr{Float64}, Ref{Int64}),
                   dh.ctx, n, last - n + 1, Δt, vals, nan_step)
        check(dh, st; t=nan_step[] * Δt, step=nan_step[])
        freq > 0 && last % freq == 0 && export_fields(dh, options, last, last * Δt)
        n = last + 1
    end
    finish_exports(dh, options)
    return dh
end

function solve!(dh::MI355XBodyDataHandler, nr::NewtonKrylov, options::AbstractJobOptions;
                precond::Symbol=:none, rel_tol::Float64=0.0)
    body = dh.body
    # displacement_bc! -> constrained dofs with their full-load values f(X)
    pts, dims, vals = Int32[], Int32[], Float64[]
    bext = [bc for bc in body.pos_single_dim_bcs if bc.field === :b_ext]
    for bc in body.pos_single_dim_bcs
        bc.field === :displacement || continue
        for i in body.point_sets[bc.point_set]
            push!(pts, i - 1); push!(dims, Int32(bc.dim) - 1)
            push!(vals, bc.fun(view(body.position, :, i)))
        end
    end
    check(dh, ccall((:pdx_set_displacement_bcs, libpdx), Int32,
                    (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Int32}, Ptr{Float64}),
                    dh.ctx, length(pts), pts, dims, vals))
    pre = Dict(:none => 0, :jacobi => 1, :learned => 2, :two_level => 3)[precond]
    opts = Ref(PdxNkOpts(nr.maxiter, nr.gmres_maxiter, nr.gmres_restart, pre, nr.tol, rel_tol,
                         nr.gmres_reltol, nr.gmres_abstol, nr.perturbation_scale))
    for n in 1:nr.n_steps
        β = n / nr.n_steps
        if !isempty(bext)
            b = zeros(3, dh.n_points)
            for bc in bext, i in body.point_sets[bc.point_set]
                v = β * bc.fun(view(body.position, :, i))
                isfinite(v) && (b[bc.dim, i] = v)
            end
            set_field!(dh, :b_ext, b)
        end
        stats = PdxNkStats()
        st = ccall((:pdx_nk_step, libpdx), Int32,
                   (Ptr{Cvoid}, Int64, Int64, Ref{PdxNkOpts}, Ref{PdxNkStats}),
                   dh.ctx, n, nr.n_steps, opts, stats)
        st == PDX_ERR_NOT_CONVERGED &&
            error("Newton-Krylov solver did not converge after $(nr.maxiter) iterations!")
        check(dh, st)
        export_fields(dh, options, n, n * nr.Δt)
    end
    finish_exports(dh, options)
    return dh
end

function init_time_solver!(vv::VelocityVerlet, dh::MI355XBodyDataHandler)
    if vv.Δt < 0
        vv.Δt = calc_stable_timestep(dh, vv.safety_factor)
    end
    if vv.end_time < 0
        vv.end_time = vv.n_steps * vv.Δt
    elseif vv.n_steps < 0
        vv.n_steps = vv.end_time ÷ vv.Δt + 1
    end
    return nothing
end

init_time_solver!(::NewtonKrylov, ::MI355XBodyDataHandler) = nothing
init_time_solver!(::DynamicRelaxation, ::MI355XBodyDataHandler) = nothing

"""
    submit_mi355x(job::Job; device=0, precision=:fp64)

Runs `job` with the body on GPU `device`; returns the data handler.
"""
function submit_mi355x(job::Job; device::Int=0, precision::Symbol=:fp64)
    dh = MI355XBodyDataHandler(job.spatial_setup; device, precision)
    init_time_solver!(job.time_solver, dh)
    initialize!(dh, job.time_solver)
    solve!(dh, job.time_solver, job.options)
    return dh
end

end # module
