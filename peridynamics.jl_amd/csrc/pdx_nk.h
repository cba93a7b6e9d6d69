// Newton-Krylov workspace and the hooks of the learned preconditioner.
#pragma once

#include "pdx_internal.h"

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
    // two-level preconditioner: aggregates of `agg` consecutive points, mc = 3 nagg
    int agg = 0;
    int64_t nagg = 0, mc = 0;
    double *Ac = nullptr;       // mc x mc, the inverse of the Galerkin operator once built
    double *rc = nullptr, *zc = nullptr;
    bool coarse_ready = false;
    int64_t n_jv = 0;
};

// Jv on device vectors (internal order)
pdx_status pdx_nk_jv_dev(pdx_ctx *c, const double *v, double *jv, double scale);

// learned preconditioner (pdx_gnn.hip)
pdx_status pdx_gnn_prepare(pdx_ctx *c);                                  // per Newton iteration
pdx_status pdx_gnn_apply(pdx_ctx *c, const double *v, double *z);        // z = M^-1 v
pdx_status pdx_gnn_record(pdx_ctx *c, const double *rhs, const double *du);  // Newton pair
pdx_status pdx_gnn_record_krylov(pdx_ctx *c, const double *z, const double *jz);  // Krylov pair
bool pdx_gnn_collecting(pdx_ctx *c);
void pdx_gnn_disable(pdx_ctx *c);  // the learned preconditioner falls back to two-level
bool pdx_gnn_active(pdx_ctx *c);   // the network is applied (trained, validated, not disabled)
pdx_status pdx_nk_build_jacobi(pdx_ctx *c);  // point-block Jacobi at the current state
pdx_status pdx_nk_build_coarse(pdx_ctx *c);  // two-level coarse operator at the current state
pdx_status pdx_nk_coarse_add(pdx_ctx *c, const double *v, double *z, double alpha);  // z += a P Ac^-1 P^T v

// preconditioners of pdx_nk_opts.precond
#define PDX_PRE_NONE 0
#define PDX_PRE_JACOBI 1
#define PDX_PRE_LEARNED 2
#define PDX_PRE_TWO_LEVEL 3
