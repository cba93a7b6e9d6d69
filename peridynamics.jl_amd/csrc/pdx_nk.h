// Newton-Krylov workspace and the hooks of the learned preconditioner.
#pragma once

#include <rocblas/rocblas.h>

#include "pdx_internal.h"

// Krylov columns the FGMRES loop keeps queued beyond the one whose convergence flag the host
// reads: after queuing column j it spins on the flag word of column j - PDX_NK_LAG (no event
// in the stream), so the device always has a column queued; PDX_NK_SLOTS flag slots rotate
// over the columns in flight.  One column suffices: two measured the same
// time per iteration and queue one more column past convergence.
#ifndef PDX_NK_LAG
#define PDX_NK_LAG 1
#endif
#define PDX_NK_SLOTS (PDX_NK_LAG + 1)


struct pdx_nk {
    int64_t m = 0;              // 3n
    int64_t nc = 0;             // constrained dofs
    int64_t *cdof = nullptr;    // internal dof index of each constraint
    double *cval = nullptr;     // full-load value of each constraint
    uint8_t *cmask = nullptr;   // 3n, 1 = constrained
    double *r = nullptr, *du = nullptr, *w = nullptr, *t0 = nullptr;  // 3n
    double *Vb = nullptr;       // (restart + 1) x 3n Krylov basis
    double *Zb = nullptr;       // restart x 3n preconditioned directions (flexible GMRES)
    int m_cap = 0;
    double4 *recp = nullptr, *recm = nullptr;  // perturbed positions
    double *fp = nullptr, *fm = nullptr;       // b_int V at the perturbed positions
    double *red = nullptr;      // reduction partials
    double *h_dev = nullptr;    // small coefficient vectors
    double *dinv = nullptr;     // 9n point-block Jacobi inverse
    double *dsc = nullptr;      // device scalars (||u||^2, ||v||^2)
    double *gm = nullptr;       // GMRES state on the device: Hessenberg, rotations, g, y, flags
    int *flag_host = nullptr;   // pinned (coherent): the convergence flags of the columns in flight
    int *flag_dev = nullptr;    // the same memory as the kernels address it
    bool jv_fused = false;      // J v by the one-kernel central difference (k_bb_jv)
    double *vparts = nullptr;   // partial sums of |v|^2 for a J v whose caller has none
    double *zparts = nullptr;   // partial sums of |z|^2 left by the preconditioner's kernels
    // two-level preconditioner: aggregates of `agg` consecutive points, nf basis fields per
    // aggregate, mc = nf nagg coarse dofs
    int agg = 0;
    int64_t nagg = 0, mc = 0;
    int nf = 0, nf_cap = 0;
    int maxs = 0;                // neighbouring aggregates of an aggregate (incl. itself)
    int32_t *aslot = nullptr;    // nagg x maxs neighbour aggregates, -1 padded
    int32_t *pJ = nullptr;       // n x NK_TP aggregates reached by each point's bonds
    int *dflag = nullptr;        // overflow flag and the rocSOLVER infos
    double *phi = nullptr;       // n x nf x 3 basis fields, orthonormal per aggregate
    double *phis = nullptr;      // the same by 64-point chunk, field, component and lane
    double *Ycs = nullptr;       // n x NK_TP x nf x 3 per-point stiffness products
    double *Ac = nullptr;        // mc x mc Galerkin operator, then its Cholesky inverse
    float *Ainv = nullptr;       // mc x mc inverse of the Galerkin operator (float32)
    double *rc = nullptr, *zc = nullptr;
    rocblas_handle rb = nullptr;
    bool coarse_ready = false;
    int64_t n_jv = 0;
    int col_tok = 0;             // token of the last queued Krylov column (1 .. 2^30 - 1)
    int tok[PDX_NK_SLOTS] = {};  // tokens of the columns in the flag slots
    bool verbose = false;       // PDX_NK_VERBOSE: a line per Newton iteration on stderr
};

// aggregate size of the two-level preconditioner: a power of two >= 64, at most 256
// aggregates (pdx_nk_aggregates doubles it where the bonds of a point would reach too many)
inline int pdx_nk_agg_size(int64_t n)
{
    int64_t agg = 64;
    while ((n + agg - 1) / agg > 256) agg *= 2;
    return (int)agg;
}

// Jv on device vectors (internal order); `parts` (np partial sums of |v|^2) when known
pdx_status pdx_nk_jv_dev(pdx_ctx *c, const double *v, double *jv, double scale,
                         const double *parts = nullptr, int np = 0);

// learned preconditioner (pdx_gnn.hip)
pdx_status pdx_gnn_prepare(pdx_ctx *c);                                  // trained check
pdx_status pdx_gnn_apply(pdx_ctx *c, const double *v, double *z);        // z = M^-1 v
// the network's coarse fields (n_fields per point, 3 floats each, internal order)
pdx_status pdx_gnn_fields(pdx_ctx *c, const float **fields, int *n_fields);
pdx_status pdx_gnn_record(pdx_ctx *c, const double *rhs, const double *du);  // Newton pair
pdx_status pdx_gnn_record_krylov(pdx_ctx *c, const double *z, const double *jz);  // Krylov pair -> error sample
bool pdx_gnn_collecting(pdx_ctx *c);
void pdx_gnn_disable(pdx_ctx *c);  // the learned preconditioner falls back to the translations
bool pdx_gnn_active(pdx_ctx *c);   // the network is applied (trained, validated, not disabled)
pdx_status pdx_nk_build_jacobi(pdx_ctx *c);  // point-block Jacobi at the current state
pdx_status pdx_nk_aggregates(pdx_ctx *c);    // sets agg and nagg once per context
pdx_status pdx_nk_build_coarse(pdx_ctx *c);  // two-level coarse operator at the current state
// the same with the network's fields appended to the translations (n_fields per point)
pdx_status pdx_nk_build_coarse_fields(pdx_ctx *c, const float *fields, int n_fields);
pdx_status pdx_nk_coarse_add(pdx_ctx *c, const double *v, double *z, double alpha);  // z += a P Ac^-1 P^T v
// z = D^-1 v + P Ac^-1 P^T v; parts (optional): the ceil(n / 256) partial sums of |z|^2
pdx_status pdx_nk_two_level(pdx_ctx *c, const double *v, double *z, double *parts = nullptr);

// preconditioners of pdx_nk_opts.precond
#define PDX_PRE_NONE 0
#define PDX_PRE_JACOBI 1
#define PDX_PRE_LEARNED 2
#define PDX_PRE_TWO_LEVEL 3
