/*
 * pdx.h -- C ABI of the MI355X force-density library (libpdx.so).
 *
 * The library replaces the force-density hot path of Peridynamics.jl for one body on one
 * GPU: the bond system (Morton-ordered points, full CSR bond lists), the bond-based (BB),
 * ordinary state-based (OSB) and correspondence (CPD) force densities, bond failure and
 * damage fused with the Velocity Verlet update, and the Jacobian-free Newton-Krylov
 * J*v products with a device-resident flexible GMRES.
 *
 * Conventions
 *  - Every host array is copied; the library never keeps a caller pointer.
 *  - Point vectors are Julia's column-major 3 x n (x, y, z of a point contiguous).
 *  - Indices are 0-based.  The Morton permutation is internal: every get/set uses the
 *    caller's point order.
 *  - Status: 0 = PDX_OK, negative = error; pdx_last_error() explains the last one.
 *  - A context is used by one host thread at a time; it owns one HIP stream and all of
 *    its device memory until pdx_destroy().
 *
 * Reference interfaces replaced (paths relative to the Peridynamics.jl repository):
 *  see the comment above each entry point.
 */
#ifndef PDX_H
#define PDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pdx_ctx pdx_ctx;
typedef int32_t pdx_status;

#define PDX_OK 0
#define PDX_ERR_ARG (-1)             /* ArgumentError */
#define PDX_ERR_HIP (-2)             /* HIP runtime failure */
#define PDX_ERR_OOM (-3)             /* device allocation failed */
#define PDX_ERR_NAN (-4)             /* NaNError(t, step), src/auxiliary/errors.jl:98-114 */
#define PDX_ERR_STATE (-5)           /* call out of order */
#define PDX_ERR_DUPLICATE_POINT (-6) /* check_point_duplicates, bond_system.jl:125-132 */
#define PDX_ERR_NOT_CONVERGED (-7)   /* newton_krylov.jl:318-324 */
#define PDX_ERR_UNSUPPORTED (-8)     /* a feature this build does not provide */

/* precision */
#define PDX_FP64 0  /* all arithmetic in FP64 */
#define PDX_MIXED 1 /* FP32 stretch/force arithmetic from FP64 differences, FP64 sums and
                       FP64 integration; BB, OSB and CPD materials (OSB with several
                       parameter sets runs the FP64 kernels); Newton-Krylov residuals and
                       Jacobian products use the FP64 kernels in both precisions */

/* material kinds */
#define PDX_BB 0  /* BBMaterial,  src/physics/bond_based.jl:74-84 */
#define PDX_OSB 1 /* OSBMaterial, src/physics/ordinary_state_based.jl:82-94 */
#define PDX_CPD 2 /* CMaterial,   src/physics/correspondence.jl:111-128 */

/* constitutive models of CMaterial (src/physics/constitutive_models.jl) */
#define PDX_SVK 0 /* SaintVenantKirchhoff :33-48 */
#define PDX_NHP 1 /* NeoHookePenalty :231-255 */
#define PDX_LINEAR_ELASTIC 2 /* LinearElastic :81-97 (Hooke matrix in Voigt form) */
#define PDX_NEO_HOOKE 3      /* NeoHooke :167-175 */

/* zero-energy-mode stabilisation of CMaterial (src/discretization/zem_stabilization.jl) */
#define PDX_ZEM_SILLING 0 /* ZEMSilling(Cs) :16-21, tzem = w beta z, beta = Cs bc / delta */
#define PDX_ZEM_WAN 1     /* ZEMWan() :30-42, tzem = w C_1 z, C_1 = C : K^-1 (correspondence.jl:253-285) */

/* fields for pdx_set_field / pdx_get_field (storage field names of the reference) */
#define PDX_FIELD_POSITION 0      /* 3 x n */
#define PDX_FIELD_DISPLACEMENT 1  /* 3 x n */
#define PDX_FIELD_VELOCITY 2      /* 3 x n */
#define PDX_FIELD_VELOCITY_HALF 3 /* 3 x n */
#define PDX_FIELD_ACCELERATION 4  /* 3 x n */
#define PDX_FIELD_B_INT 5         /* 3 x n */
#define PDX_FIELD_B_EXT 6         /* 3 x n */
#define PDX_FIELD_DAMAGE 7        /* n */
#define PDX_FIELD_N_ACTIVE_BONDS 8 /* n, returned as double */
#define PDX_FIELD_DEFGRAD 9       /* 9 x n, column-major 3x3 per point (CPD) */
#define PDX_FIELD_CAUCHY_STRESS 10 /* 9 x n (CPD) */
#define PDX_FIELD_REF_POSITION 11 /* 3 x n, read only */
#define PDX_FIELD_VOLUME 12       /* n, read only */
#define PDX_FIELD_VELOCITY_HALF_OLD 13 /* 3 x n, DynamicRelaxation */
#define PDX_FIELD_B_INT_OLD 14    /* 3 x n, DynamicRelaxation */
#define PDX_FIELD_DENSITY_MATRIX 15 /* n, read only: the DynamicRelaxation Lambda of a point
                                     (the reference stores it 3 x n with equal rows) */

/* Influence functions omega(delta, L) (src/discretization/kernels.jl). */
#define PDX_KERNEL_LINEAR 0             /* linear_kernel, delta / L (:22), the default */
#define PDX_KERNEL_CONST_ONE 1          /* const_one_kernel, 1 (:10) */
#define PDX_KERNEL_CUBIC_BSPLINE 2      /* cubic_b_spline_kernel (:43-52) */
#define PDX_KERNEL_CUBIC_BSPLINE_NORM 3 /* cubic_b_spline_kernel_norm (:74-84) */

/* Point parameters of one material (StandardPointParameters, src/core/parameters.jl:107-126;
 * CPointParameters, src/physics/correspondence.jl:148-169).  The host derives them as
 * material! does (src/physics/material_parameters.jl:64-231, src/physics/fracture.jl:110-130,
 * bc = 18K/(pi delta^4) at src/physics/bond_based.jl:89). */
typedef struct pdx_params {
    double delta, rho, E, nu, G, K, lambda, mu, Gc, epsc, bc;
    int32_t model;  /* PDX_SVK, PDX_NHP, PDX_LINEAR_ELASTIC or PDX_NEO_HOOKE (CPD only) */
    int32_t kernel; /* influence function of OSB and CPD (PDX_KERNEL_*); BB has none */
    double zem_cs;  /* ZEMSilling Cs, src/discretization/zem_stabilization.jl:16-21 */
    double maxdmg;  /* CMaterial maxdmg, correspondence.jl:123-128 */
    int32_t zem;    /* PDX_ZEM_SILLING or PDX_ZEM_WAN (CPD only) */
    int32_t reserved_;
} pdx_params;

/* Options and statistics of one Newton-Krylov load step (NewtonKrylov,
 * src/time_solvers/newton_krylov.jl:37-127). */
typedef struct pdx_nk_opts {
    int32_t maxiter;          /* Newton iterations */
    int32_t gmres_maxiter;    /* total Krylov iterations per linear solve */
    int32_t gmres_restart;    /* restart length */
    int32_t precond;          /* 0 none (the reference), 1 point-block Jacobi, 2 learned
                                 (two-level whose coarse space adds the message-passing
                                 network's fields), 3 two-level (Jacobi + Galerkin coarse
                                 correction by the translations of point aggregates) */
    double tol;               /* absolute tolerance on ||r|| (reference semantics) */
    double rel_tol;           /* > 0: stop at ||r|| < rel_tol * ||r_0|| of the load step */
    double gmres_reltol, gmres_abstol;
    double perturbation_scale;
} pdx_nk_opts;

typedef struct pdx_nk_stats {
    int32_t newton_iters;
    int32_t krylov_iters;
    int32_t jv_evals;
    int32_t precond_fallbacks; /* learned solves that missed the tolerance without reducing the
                                  linear residual tenfold, redone with the two-level
                                  preconditioner of the translations */
    double residual_first, residual_final;
    double seconds;
} pdx_nk_stats;

/* --- context -------------------------------------------------------------------------- */
const char *pdx_version(void);
/* Creates a context on HIP device `device`.  precision = PDX_FP64 or PDX_MIXED. */
pdx_status pdx_create(pdx_ctx **out, int32_t device, int32_t precision);
void pdx_destroy(pdx_ctx *ctx);
const char *pdx_last_error(const pdx_ctx *ctx);
pdx_status pdx_synchronize(pdx_ctx *ctx);

/* --- body ----------------------------------------------------------------------------- */
/* Body positions and volumes (Body, src/discretization/body.jl:83-125) and the per-point
 * failure permission (failure_permit!, src/physics/fracture.jl:38-42). */
pdx_status pdx_set_points(pdx_ctx *ctx, int64_t n, const double *X, const double *volume,
                          const uint8_t *fail_permit);
/* One parameter set for the whole body (material!, material_parameters.jl:64-92). */
pdx_status pdx_set_material(pdx_ctx *ctx, int32_t kind, const pdx_params *params);
/* Several parameter sets on one body (material!(body, set_name; ...) called per point set,
 * material_parameters.jl:64-92, kept as point_params + params_map; ParameterHandler,
 * src/core/parameter_handler.jl:36-81).  point_set[i] in [0, n_sets) is the set of caller
 * point i (0-based params_map).  Kinds PDX_BB, PDX_OSB and PDX_CPD; all sets share one
 * horizon (a CMaterial's sets also its model, zem_cs and maxdmg); at most 255 sets.  A bond
 * uses the mean of both ends' bc (BB) or G and K (OSB) and the owner's eps_c
 * (bond_based.jl:163-180, ordinary_state_based.jl:151-174); a CMaterial point its own set
 * (correspondence.jl:183-191).  n_sets = 1 is pdx_set_material.  Call after
 * pdx_set_points and before pdx_build_bonds. */
pdx_status pdx_set_material_sets(pdx_ctx *ctx, int32_t kind, int32_t n_sets, const pdx_params *sets,
                                 const int32_t *point_set);
/* Surface correction of the material's type parameter (BBMaterial{Correction},
 * OSBMaterial{Correction}; src/discretization/bond_system_corrections.jl): 0 NoCorrection,
 * 1 EnergySurfaceCorrection (:58-61).  With 1, pdx_build_bonds computes mfactor and the
 * per-bond scfactor as initialize! does (:76-96, calc_mfactor! :98-122, calc_scfactor!
 * :161-204) and every step multiplies each bond's weight by its scfactor.  PDX_BB and
 * PDX_OSB.  Call before pdx_build_bonds. */
pdx_status pdx_set_surface_correction(pdx_ctx *ctx, int32_t kind);
/* The correction's factors: scfactor[n_bonds] in the bond order of pdx_get_bonds, and
 * mfactor[n][3] (system.correction.scfactor / .mfactor).  Either pointer may be null. */
pdx_status pdx_get_surface_correction(pdx_ctx *ctx, double *scfactor, double *mfactor);
/* precrack!(body, set_a, set_b; update_dmg), src/discretization/predefined_cracks.jl:106-112:
 * filter_bonds = 1 deletes the bonds (update_dmg = false), 0 breaks them at t = 0. */
pdx_status pdx_add_precrack(pdx_ctx *ctx, int64_t na, const int32_t *set_a, int64_t nb,
                            const int32_t *set_b, int32_t filter_bonds);
/* find_bonds + precracks + bond state (BondSystem, bond_system.jl:55-132, 142-176;
 * apply_precracks!, predefined_cracks.jl:31-61).  Bonds of a point are in ascending
 * neighbour index order. */
pdx_status pdx_build_bonds(pdx_ctx *ctx, int64_t *n_bonds_out);
/* The bond list in the caller's point order: row_ptr[n+1], nbr, L, fail_permit, and the
 * bond_active state (any pointer may be NULL).  Replaces reading system.bonds /
 * storage.bond_active. */
pdx_status pdx_get_bonds(pdx_ctx *ctx, int64_t *row_ptr, int32_t *nbr, double *L,
                         uint8_t *fail_permit, uint8_t *active);

/* --- fields ----------------------------------------------------------------------------- */
pdx_status pdx_set_field(pdx_ctx *ctx, int32_t field, const double *host);
pdx_status pdx_get_field(pdx_ctx *ctx, int32_t field, double *host);

/* --- export pipeline -------------------------------------------------------------------- */
/* Device side of export_results / export_reference_results
 * (src/core/threads_body_data_handler.jl:197-212; _export_results, src/auxiliary/io.jl:157-167):
 * takes a snapshot of `fields` (PDX_FIELD_*, at most 16) in the caller's point order on
 * the compute stream and copies it to pinned host memory on a second stream; returns at
 * once, so the following steps overlap with the copy.  `tag` is returned by the matching
 * wait (e.g. the time step).  At most two exports may be pending. */
pdx_status pdx_export_begin(pdx_ctx *ctx, int32_t n_fields, const int32_t *fields, int64_t tag);
/* Waits for the oldest pending export; field_data[k] points to field k (n * components
 * doubles, point-major) in the library's pinned memory of slot *slot (0 or 1).  Exports
 * fill the two slots in turn: the data stays valid until the next pdx_export_begin that
 * fills the same slot. */
pdx_status pdx_export_wait(pdx_ctx *ctx, int64_t *tag, int32_t *slot, const double **field_data);

/* --- force density and Velocity Verlet --------------------------------------------------- */
/* calc_stable_timestep(dh, safety_factor): velocity_verlet.jl:136-163 with
 * calc_timestep_point, bond_system.jl:269-277. */
pdx_status pdx_stable_timestep(pdx_ctx *ctx, double safety_factor, double *dt_out);
/* calc_force_density!(dh, t, dt) at the current positions: failure, damage, force density
 * (threads_body_data_handler.jl:184-195, bond_system.jl:285-297) plus the NaN check of
 * nans.jl:10-18 (returns PDX_ERR_NAN). */
pdx_status pdx_force_density(pdx_ctx *ctx, double t, double dt);
/* Velocity boundary conditions on velocity_half (velocity_bc!, boundary_conditions.jl:
 * 422-435).  Point k of the list is points[k]; src[3k+d] is the index into the per-step
 * value vector that sets its dimension d, or -1.  The host evaluates the reference's
 * closures (SingleDimBC f(t): one value per condition; PosDepSingleDimBC f(p, t): one
 * value per point) and a non-finite value leaves the dof alone (apply_sdbc!, apply_pdsdbc!).
 * Replaces the BC part of apply_boundary_conditions!, boundary_conditions.jl:212-226. */
pdx_status pdx_set_velocity_bcs(pdx_ctx *ctx, int64_t n_points, const int32_t *points,
                                const int32_t *src, int64_t n_src);
/* Time-dependent force-density conditions on b_ext (forcedensity_bc! with f(t) or f(p, t),
 * boundary_conditions.jl:559-575), applied at the start of every step as
 * apply_boundary_conditions! does.  src indexes the per-step value rows whose length
 * pdx_set_velocity_bcs set (call that first, with the total row length). */
pdx_status pdx_set_force_density_bcs(pdx_ctx *ctx, int64_t n_points, const int32_t *points,
                                     const int32_t *src);
/* n_steps Velocity Verlet steps n = first_step ... first_step + n_steps - 1 with step size
 * dt (solve!, verlet_timestep!, update_vel_half!, update_disp_and_pos!,
 * _update_acc_and_vel!: velocity_verlet.jl:165-295).  bc_values holds n_steps rows of
 * n_src values, row k for time (first_step + k) * dt.  One fused launch per step.  On a
 * NaN force density returns PDX_ERR_NAN and the first such step in *nan_step_out. */
pdx_status pdx_vv_run(pdx_ctx *ctx, int64_t first_step, int64_t n_steps, double dt,
                      const double *bc_values, int64_t *nan_step_out);
/* pdx_vv_run without any device allocation (no counterpart in the reference).  For callers
 * that cannot allocate device memory while they run -- tight stepping or serving loops,
 * harnesses that forbid hipMalloc -- once the context is set up: pdx_vv_run may hipFree and
 * hipMalloc its boundary-condition value table when a run needs a larger one, this variant
 * stages the table in a fixed device array of PDX_NOALLOC_BC_VALUES doubles (4 MiB) that
 * each GPU holds from the first call until the process ends.  A run with n_src * n_steps
 * above that returns PDX_ERR_ARG (split it into shorter runs).  The array is shared by all
 * contexts on a GPU: calls that overlap on one GPU race.  Results equal pdx_vv_run's. */
#define PDX_NOALLOC_BC_VALUES (1 << 19)
pdx_status pdx_vv_run_noalloc(pdx_ctx *ctx, int64_t first_step, int64_t n_steps, double dt,
                              const double *bc_values, int64_t *nan_step_out);

/* --- dynamic relaxation ----------------------------------------------------------------- */
/* DynamicRelaxation set-up (init_density_matrix!, dynamic_relaxation.jl:119-176): the
 * per-point Lambda = damping_factor / 4 dt^2 n_bonds 5 pi delta^2 bc; zeroes
 * velocity_half_old and b_int_old on the first call.  Bond-based bodies. */
pdx_status pdx_dr_init(pdx_ctx *ctx, double dt, double damping_factor);
/* Steps first .. first + n_steps - 1 of solve!(dh, ::DynamicRelaxation) (relaxation_timestep!,
 * relaxation_first_step!, relaxation_step!, relaxation_updates!: dynamic_relaxation.jl:
 * 219-350).  bc_values as for pdx_vv_run (velocity_half conditions, row k for step first+k). */
pdx_status pdx_dr_run(pdx_ctx *ctx, int64_t first_step, int64_t n_steps, double dt,
                      const double *bc_values, int64_t *nan_step_out);
/* calc_damping (dynamic_relaxation.jl:262-323) of the current fields. */
pdx_status pdx_dr_damping(pdx_ctx *ctx, double dt, double *cn_out);

/* --- Newton-Krylov ---------------------------------------------------------------------- */
/* Constrained dofs from displacement_bc! (PosSingleDimBC, boundary_conditions.jl:130-177,
 * 675-688): dof 3*points[k] + dims[k] is prescribed with the full-load value values[k]
 * (the host evaluates f(p)); load step factor beta scales it. */
pdx_status pdx_set_displacement_bcs(pdx_ctx *ctx, int64_t n, const int32_t *points,
                                    const int32_t *dims, const double *values);
/* Jacobian-vector product of the residual r(u) = (b_int + b_ext) V with constrained rows
 * set to identity (jacobian_vector_product!, newton_krylov.jl:504-671): Jv = J v for the
 * current displacement.  v and Jv are 3n host vectors in the caller's order. */
pdx_status pdx_jv(pdx_ctx *ctx, const double *v, double *Jv, double perturbation_scale);
/* Residual r = (b_int + b_ext) V, constrained dofs zero (calc_residual!, :371-387). */
pdx_status pdx_residual(pdx_ctx *ctx, double *r, double *norm_out);
/* The aggregates of the two-level preconditioner (no counterpart in the reference): runs of
 * *agg consecutive internal (Morton) points, *n_aggregates of them, and the coarse fields
 * per aggregate of the last coarse space built (3 translations, + 8 learned; 0 before). */
pdx_status pdx_nk_coarse_info(pdx_ctx *ctx, int32_t *agg, int64_t *n_aggregates, int32_t *fields);
/* One load step n of n_steps (newton_krylov_step!, newton_krylov.jl:298-332): applies the
 * scaled displacement BCs, then Newton iterations with a device-resident flexible GMRES
 * (replacing IterativeSolvers.gmres!, :448-450). */
pdx_status pdx_nk_step(pdx_ctx *ctx, int64_t n, int64_t n_steps, const pdx_nk_opts *opts,
                       pdx_nk_stats *stats);

/* --- learned preconditioner (no counterpart in the reference) ---------------------------- */
/* The reference solves J du = -r with IterativeSolvers.gmres! and no preconditioner
 * (newton_krylov.jl:448-450).  These entry points build, train and expose the
 * message-passing network behind opts.precond = 2: the network maps the reference
 * configuration (positions, constraints, bond counts) to 8 vector fields per point, and the
 * learned preconditioner is the two-level preconditioner whose coarse space on every
 * aggregate spans the three translations and those 8 fields.  BBMaterial bodies of at most
 * 131072 points.
 *
 * While collection is enabled, every Krylov pair (J z, z) of a solve preconditioned by the
 * two-level preconditioner of the translations (M_T) yields the error sample
 * e = z - M_T(J z), reservoir-sampled into at most max_krylov_pairs float32 vectors on the
 * device. */
pdx_status pdx_gnn_collect(pdx_ctx *ctx, int32_t enable, int64_t max_krylov_pairs);
/* n_newton is always 0; n_krylov = the error samples held. */
pdx_status pdx_gnn_pairs(pdx_ctx *ctx, int64_t *n_newton, int64_t *n_krylov);
/* `steps` Adam steps (cosine-decayed step size lr) on random subsets of `batch` samples;
 * the last tenth of the samples is held out.  Loss per sample and aggregate I: the ridge
 * least-squares residual of e on I by [translations | fields] over |e|^2.  *loss_out = the
 * loss of the last reported step. */
pdx_status pdx_gnn_train(pdx_ctx *ctx, int32_t steps, int32_t batch, double lr, uint64_t seed,
                         double *loss_out);
/* Held-out validation: *val_loss = the held-out loss with the fields over the loss of the
 * translations alone (1 = the fields add nothing).  The learned coarse space is used while
 * that ratio is below `threshold` (0.95 after training) and until a learned solve needs the
 * fallback (pdx_nk_stats.precond_fallbacks); otherwise the preconditioner is the two-level
 * one of the translations.  threshold NaN only reports. */
pdx_status pdx_gnn_gate(pdx_ctx *ctx, double threshold, double *val_loss, int32_t *use_network);
/* Number of float32 parameters of the network. */
int64_t pdx_gnn_n_params(void);
/* Reads (theta_out) or replaces (theta_in) the parameters; replacing marks the network as
 * trained.  *out_scale is returned as 1 (the coarse solve fixes the scale of the fields). */
pdx_status pdx_gnn_params(pdx_ctx *ctx, float *theta_out, const float *theta_in,
                          double *out_scale);
/* Test access, all in the library's internal (Morton) point order: permutation
 * perm[internal] = caller index, near-bond graph (CSR), the network's 16 input features per
 * point, point-block Jacobi inverse at the current state (9 per point, column-major),
 * constrained mask (3 per point).  Any pointer may be NULL; call with only n_near first. */
pdx_status pdx_gnn_graph(pdx_ctx *ctx, int64_t *n_near, int64_t *perm, int64_t *near_ptr,
                         int32_t *near, float *features, double *dinv, uint8_t *cmask);
/* z = M(y), the learned two-level preconditioner at the current state, for one 3n vector
 * (internal order). */
pdx_status pdx_gnn_apply_vec(pdx_ctx *ctx, const double *y, double *z);
/* The network's fields for the current parameters: 24 floats per point (8 fields x 3),
 * internal order. */
pdx_status pdx_gnn_field_values(pdx_ctx *ctx, float *fields);
/* Loss and gradient of the current parameters over n_samples error samples (Y: n_samples
 * rows of 3n, internal order; X unused) without updating them.  The samples replace the
 * collected ones. */
pdx_status pdx_gnn_loss_grad(pdx_ctx *ctx, int32_t n_samples, const double *Y, const double *X,
                             double *loss, float *grad);

/* --- instrumentation ---------------------------------------------------------------------- */
/* Per-launch HIP-event timing of the step kernels (off by default). */
pdx_status pdx_set_profiling(pdx_ctx *ctx, int32_t enable);
/* Average duration (ms) and count of the timed launches of kernel `name`
 * ("step" = the fused Velocity Verlet kernel, "pass1" = first pass of OSB/CPD). */
pdx_status pdx_kernel_time(pdx_ctx *ctx, const char *name, double *avg_ms, int64_t *count);
/* Sizes: n points, n bonds, device bytes held. */
pdx_status pdx_info(pdx_ctx *ctx, int64_t *n_points, int64_t *n_bonds, int64_t *device_bytes);
/* Device layout of the bond lists (no counterpart in the reference, whose bond lists are
 * plain vectors): 1 if the bond-based step stages neighbour data in LDS, the number of bond
 * slots including the padding of 64-point slices, the total and the largest per-block
 * neighbour table.  Used to state the step kernel's algorithmic bytes. */
pdx_status pdx_layout(pdx_ctx *ctx, int32_t *lds_mode, int64_t *bond_slots,
                      int64_t *halo_entries, int32_t *halo_max);

#ifdef __cplusplus
}
#endif

#endif /* PDX_H */
