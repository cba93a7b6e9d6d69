// Jacobian-free Newton-Krylov on the device.
//
// Reference: src/time_solvers/newton_krylov.jl.  One load step (newton_krylov_step!,
// :298-332) applies the scaled displacement conditions (apply_incr_boundary_conditions! +
// update_position!, :334-357), then iterates: force density (calc_force_density!),
// residual r = (b_int + b_ext) V with constrained rows zero (calc_residual!, :371-387),
// ||r|| against the tolerance (:389-413), and the linear solve J du = -r (:415-460).  The
// reference calls IterativeSolvers.gmres! with no preconditioner; here the solve is a
// device-resident restarted flexible GMRES (right preconditioning, classical Gram-Schmidt
// with one reorthogonalisation pass, Givens rotations on the host) whose matrix-vector
// product is jacobian_vector_product! (:504-671): two force evaluations at X + (u +- eps v)
// with eps = scale cbrt(eps_mach) (||u|| + 1) / ||v|| and identity rows for constrained
// dofs.  Preconditioners: none (the reference), point-block Jacobi from the analytic
// linearisation of the bond-based force, or the learned message-passing network
// (pdx_gnn.hip) with Jacobi as fallback.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include <rocsolver/rocsolver.h>

#include "pdx_internal.h"
#include "pdx_nk.h"

#define RED_BLOCKS 512
#define PDX_MAX_RESTART 128
// device scalars of the Krylov solver
#define NK_UN2 0  // ||u||^2
#define NK_VN2 1  // ||v||^2 of the current J v product

void pdx_nk_free(pdx_ctx *c)
{
    pdx_nk *k = c->nk;
    if (!k) return;
    for (void *p : {(void *)k->cdof, (void *)k->cval, (void *)k->cmask, (void *)k->r,
                    (void *)k->du, (void *)k->w, (void *)k->t0, (void *)k->Vb, (void *)k->Zb,
                    (void *)k->recp, (void *)k->recm, (void *)k->fp, (void *)k->fm,
                    (void *)k->red, (void *)k->dinv, (void *)k->h_dev, (void *)k->dsc,
                    (void *)k->Ac, (void *)k->rc, (void *)k->zc, (void *)k->Ainv, (void *)k->phi,
                    (void *)k->Ycs, (void *)k->pJ, (void *)k->aslot, (void *)k->dflag})
        if (p) (void)hipFree(p);
    if (k->rb) rocblas_destroy_handle(k->rb);
    delete k;
    c->nk = nullptr;
}

// ----------------------------------------------------------------------------------------
// small vector kernels (3n-long vectors in internal order)
// ----------------------------------------------------------------------------------------
// partial[blk * nv + k] = sum over this block's share of V_k . w   (k < nv); eight
// vectors per pass over w, one block reduction per eight dot products
#define MD_CHUNK 8
__global__ __launch_bounds__(256) void k_multidot(int64_t m, int nv, const double *V,
                                                  int64_t ldv, const double *w, double *partial)
{
    __shared__ double sh[MD_CHUNK][256];
    for (int k0 = 0; k0 < nv; k0 += MD_CHUNK) {
        const int kn = min(MD_CHUNK, nv - k0);
        double s[MD_CHUNK];
#pragma unroll
        for (int q = 0; q < MD_CHUNK; q++) s[q] = 0.0;
        for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (int64_t)gridDim.x * 256) {
            const double wi = w[i];
#pragma unroll
            for (int q = 0; q < MD_CHUNK; q++)
                if (q < kn) s[q] += V[(int64_t)(k0 + q) * ldv + i] * wi;
        }
#pragma unroll
        for (int q = 0; q < MD_CHUNK; q++) sh[q][threadIdx.x] = s[q];
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o)
#pragma unroll
                for (int q = 0; q < MD_CHUNK; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
            __syncthreads();
        }
        if ((int)threadIdx.x < kn) partial[(int64_t)blockIdx.x * nv + k0 + threadIdx.x] = sh[threadIdx.x][0];
        __syncthreads();
    }
}

// out[k] = sum over blocks of partial[blk * nv + k]; one workgroup per k
__global__ __launch_bounds__(256) void k_reduce_partials(int nblk, int nv, const double *partial,
                                                         double *out)
{
    __shared__ double sh[256];
    const int k = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) s += partial[(int64_t)b * nv + k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = sh[0];
}

// w -= sum_k h[k] V_k
__global__ void k_multiaxpy(int64_t m, int nv, const double *V, int64_t ldv, const double *h,
                           double *w)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double s = w[i];
    for (int k = 0; k < nv; k++) s -= h[k] * V[k * ldv + i];
    w[i] = s;
}

// y = a * x
__global__ void k_scale(int64_t m, double a, const double *x, double *y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) y[i] = a * x[i];
}

// x += sum_k y[k] Z_k
__global__ void k_lincomb_add(int64_t m, int nv, const double *Z, int64_t ldz, const double *y,
                              double *x)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double s = x[i];
    for (int k = 0; k < nv; k++) s += y[k] * Z[k * ldz + i];
    x[i] = s;
}

// r = b - r_in  (used as r = b - A x)
__global__ void k_sub_from(int64_t m, const double *b, double *r)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) r[i] = b[i] - r[i];
}

// apply_incr_boundary_conditions! + update_position! (constrained dofs):
// u = beta * f(X); x = X + u; a non-finite value leaves the dof alone
__global__ void k_nk_bcs(int64_t nc, const int64_t *cdof, const double *cval, double beta,
                         double *u, double4 *rec, const double4 *Xr)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    const int64_t dof = cdof[k];
    const double val = beta * cval[k];
    if (!isfinite(val)) return;
    u[dof] = val;
    const int64_t i = dof / 3;
    const int d = (int)(dof - 3 * i);
    reinterpret_cast<double *>(rec)[4 * i + d] = reinterpret_cast<const double *>(Xr)[4 * i + d] + val;
}

// calc_residual!: r = (b_int + b_ext) V, constrained rows zero
__global__ void k_residual(int64_t n, const double *bint, const double *bext,
                           const double4 *Xr, const uint8_t *cmask, double *r)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * n) return;
    const int64_t i = t / 3;
    r[t] = cmask[t] ? 0.0 : (bint[t] + bext[t]) * Xr[i].w;
}

// perturbation of jacobian_vector_product! (compute_perturbation, newton_krylov.jl:537-567):
// s = scale cbrt(eps) (||u|| + 1) / ||v||, ||v|| < eps -> 1; the squared norms are device
// scalars so that no host round trip sits between the Krylov kernels
__device__ __forceinline__ double nk_perturbation(const double *un2, const double *vn2, double sc)
{
    double vn = sqrt(*vn2);
    if (vn < 2.220446049250313e-16) vn = 1.0;
    return sc * (sqrt(*un2) + 1.0) / vn;
}

// positions X + (u + s v) and X + (u - s v) with the volume in .w
// (store_state_and_apply_perturbation!, store_pos_force_and_apply_neg_perturbation!)
__global__ void k_perturb(int64_t n, const double4 *Xr, const double *u, const double *v,
                          const double *un2, const double *vn2, double sc, double4 *rp, double4 *rm)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = nk_perturbation(un2, vn2, sc);
    const double4 X = Xr[i];
    double up[3], um[3];
    for (int d = 0; d < 3; d++) {
        up[d] = u[3 * i + d] + s * v[3 * i + d];
        um[d] = u[3 * i + d] - s * v[3 * i + d];
    }
    rp[i] = make_double4(X.x + up[0], X.y + up[1], X.z + up[2], X.w);
    rm[i] = make_double4(X.x + um[0], X.y + um[1], X.z + um[2], X.w);
}

// compute_jvp_and_restore!: Jv = (F+ V - F- V) / (2 s), constrained rows Jv = v
__global__ void k_jv_combine(int64_t m, const double *fp, const double *fm, const double *un2,
                             const double *vn2, double sc, const uint8_t *cmask, const double *v,
                             double *jv)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const double s = nk_perturbation(un2, vn2, sc);
    jv[t] = cmask[t] ? v[t] : (fp[t] - fm[t]) / (2.0 * s);
}

// x *= 1 / sqrt(*n2) (no-op for a zero norm)
__global__ void k_scale_inv_sqrt(int64_t m, const double *n2, const double *x, double *y)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const double nn = *n2;
    y[t] = nn > 0.0 ? x[t] / sqrt(nn) : x[t];
}

// update_displacement_from_du!: free dofs u += du, x += du
__global__ void k_nk_update(int64_t n, const double *du, const uint8_t *cmask, double *u,
                            double4 *rec)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * n) return;
    if (cmask[t]) return;
    const int64_t i = t / 3;
    const int d = (int)(t - 3 * i);
    u[t] = u[t] + du[t];
    double *x = reinterpret_cast<double *>(rec) + 4 * i + d;
    *x = *x + du[t];
}

// Point-block Jacobi of the bond-based residual: J_ii = -V_i sum_j bc V_j
// [ (1/L) n n^T + (eps/l) (I - n n^T) ] with n = dx/l (the derivative of r_i = b_i V_i with
// respect to u_i).  Constrained dofs get identity rows/columns.  Stores the inverse.
__global__ __launch_bounds__(256) void k_bb_block_jacobi(int64_t n, const int64_t *slice_ptr,
                                                         const int32_t *cnt, const uint32_t *bw,
                                                         const double *L, const double4 *rec,
                                                         double bc, const pdx_pset *ps,
                                                         const uint8_t *pid, const uint8_t *cmask,
                                                         double *dinv)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t b0 = slice_ptr[i >> 6] + (i & 63);
    const double4 ri = rec[i];
    double D[6] = {0, 0, 0, 0, 0, 0};  // xx yx zx yy zy zz
    for (int t = 0; t < cnt[i]; t++) {
        const int64_t b = b0 + (int64_t)t * 64;
        const uint32_t wd = bw[b];
        if (!(wd & PDX_BW_ACT)) continue;
        const double4 rj = rec[wd & PDX_BW_NBR];
        const double dx = rj.x - ri.x, dy = rj.y - ri.y, dz = rj.z - ri.z;
        const double l = sqrt((dx * dx + dy * dy) + dz * dz);
        const double nx = dx / l, ny = dy / l, nz = dz / l;
        const double Lb = L[b];
        const double eps = (l - Lb) / Lb;
        const double bcb = ps ? (ps[pid[i]].bc + ps[pid[wd & PDX_BW_NBR]].bc) / 2.0 : bc;
        const double a = bcb * rj.w / Lb, g = bcb * rj.w * eps / l;
        D[0] += (a - g) * nx * nx + g;
        D[1] += (a - g) * ny * nx;
        D[2] += (a - g) * nz * nx;
        D[3] += (a - g) * ny * ny + g;
        D[4] += (a - g) * nz * ny;
        D[5] += (a - g) * nz * nz + g;
    }
    double M[9] = {-D[0] * ri.w, -D[1] * ri.w, -D[2] * ri.w, -D[1] * ri.w, -D[3] * ri.w,
                   -D[4] * ri.w, -D[2] * ri.w, -D[4] * ri.w, -D[5] * ri.w};
    for (int d = 0; d < 3; d++) {
        if (cmask[3 * i + d]) {
            for (int e = 0; e < 3; e++) {
                M[d + 3 * e] = 0.0;
                M[e + 3 * d] = 0.0;
            }
            M[d + 3 * d] = 1.0;
        }
    }
    // inverse by the adjugate
    const double c00 = M[4] * M[8] - M[7] * M[5], c01 = M[7] * M[2] - M[1] * M[8],
                 c02 = M[1] * M[5] - M[4] * M[2];
    double det = M[0] * c00 + M[3] * c01 + M[6] * c02;
    double R[9];
    if (!(fabs(det) > 0) || !isfinite(det)) {
        for (int k = 0; k < 9; k++) R[k] = (k % 4 == 0) ? 1.0 : 0.0;
    } else {
        const double id = 1.0 / det;
        R[0] = c00 * id;
        R[1] = c01 * id;
        R[2] = c02 * id;
        R[3] = (M[6] * M[5] - M[3] * M[8]) * id;
        R[4] = (M[0] * M[8] - M[6] * M[2]) * id;
        R[5] = (M[3] * M[2] - M[0] * M[5]) * id;
        R[6] = (M[3] * M[7] - M[6] * M[4]) * id;
        R[7] = (M[6] * M[1] - M[0] * M[7]) * id;
        R[8] = (M[0] * M[4] - M[3] * M[1]) * id;
    }
    for (int k = 0; k < 9; k++) dinv[9 * i + k] = R[k];
}

// z = Dinv v (3x3 per point, column-major)
__global__ void k_apply_block(int64_t n, const double *dinv, const double *v, double *z)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *R = dinv + 9 * i;
    const double a = v[3 * i], b = v[3 * i + 1], cc = v[3 * i + 2];
    z[3 * i] = R[0] * a + R[3] * b + R[6] * cc;
    z[3 * i + 1] = R[1] * a + R[4] * b + R[7] * cc;
    z[3 * i + 2] = R[2] * a + R[5] * b + R[8] * cc;
}

// ----------------------------------------------------------------------------------------
// host helpers
// ----------------------------------------------------------------------------------------
static inline dim3 grid1(int64_t m, int bs = 256) { return dim3((unsigned)((m + bs - 1) / bs)); }

static pdx_status nk_ensure(pdx_ctx *c, int restart)
{
    if (!c->nk) {
        c->nk = new pdx_nk();
        // PDX_NK_VERBOSE=1: one line per Newton iteration on stderr (diagnostics)
        const char *e = getenv("PDX_NK_VERBOSE");
        c->nk->verbose = e && atoi(e) > 0;
    }
    pdx_nk *k = c->nk;
    const int64_t n = c->n, m = 3 * n;
    k->m = m;
    if (!k->r) {
        PDX_TRY(pdx_alloc(c, &k->cmask, m));
        PDX_HIP(c, hipMemset(k->cmask, 0, m));
        for (double **p : {&k->r, &k->du, &k->w, &k->t0, &k->fp, &k->fm})
            PDX_TRY(pdx_alloc(c, p, m));
        PDX_TRY(pdx_alloc(c, &k->recp, n));
        PDX_TRY(pdx_alloc(c, &k->recm, n));
        PDX_TRY(pdx_alloc(c, &k->red, (size_t)RED_BLOCKS * (PDX_MAX_RESTART + 2)));
        PDX_TRY(pdx_alloc(c, &k->h_dev, 2 * (PDX_MAX_RESTART + 1) + 8));
        PDX_TRY(pdx_alloc(c, &k->dsc, 8));
        PDX_HIP(c, hipMemset(k->dsc, 0, sizeof(double) * 8));
        PDX_TRY(pdx_alloc(c, &k->dinv, 9 * n));
    }
    if (restart > k->m_cap) {
        if (k->Vb) (void)hipFree(k->Vb);
        if (k->Zb) (void)hipFree(k->Zb);
        k->Vb = k->Zb = nullptr;
        PDX_TRY(pdx_alloc(c, &k->Vb, (size_t)(restart + 1) * m));
        PDX_TRY(pdx_alloc(c, &k->Zb, (size_t)restart * m));
        k->m_cap = restart;
    }
    return PDX_OK;
}

// out_dev[k] = V_k . w for k < nv (V with leading dimension m), result in device memory
static pdx_status multidot_dev(pdx_ctx *c, int nv, const double *V, const double *w, double *out_dev)
{
    pdx_nk *k = c->nk;
    if (nv > PDX_MAX_RESTART + 1) return pdx_fail(c, PDX_ERR_UNSUPPORTED, "restart too long");
    hipLaunchKernelGGL(k_multidot, dim3(RED_BLOCKS), dim3(256), 0, c->stream, k->m, nv, V, k->m, w,
                       k->red);
    PDX_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)nv), dim3(256), 0, c->stream, RED_BLOCKS, nv,
                       k->red, out_dev);
    PDX_HIP(c, hipGetLastError());
    return PDX_OK;
}

// dots on the host (one synchronisation)
static pdx_status multidot(pdx_ctx *c, int nv, const double *V, const double *w, double *dots)
{
    pdx_nk *k = c->nk;
    PDX_TRY(multidot_dev(c, nv, V, w, k->h_dev));
    PDX_HIP(c, hipMemcpyAsync(dots, k->h_dev, sizeof(double) * nv, hipMemcpyDeviceToHost, c->stream));
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    return PDX_OK;
}

static pdx_status norm2(pdx_ctx *c, const double *x, double *out)
{
    double d;
    PDX_TRY(multidot(c, 1, x, x, &d));
    *out = std::sqrt(d);
    return PDX_OK;
}

// force density * V at the positions `rec_in` into `out` (calc_force_density! at a
// perturbed state, b_int * volume)
static pdx_status force_scaled(pdx_ctx *c, const double4 *rec_in, double *out)
{
    pdx_step_args a;
    // finite-difference products need the FP64 force also in a mixed-precision context
    PDX_TRY(pdx_fill_step_args(c, &a, 0, PDX_F_FORCE_ONLY | PDX_F_SCALE_VOL | PDX_F_FP64, 0.0, nullptr));
    a.rec_in = rec_in;
    a.bint_out = out;
    return pdx_launch_step(c, a);
}

// Jv for the current displacement (jacobian_vector_product!).  ||u||^2 is the device
// scalar NK_UN2, kept current by nk_refresh_unorm(); ||v||^2 is formed here on the device.
static pdx_status nk_refresh_unorm(pdx_ctx *c)
{
    return multidot_dev(c, 1, c->u, c->u, c->nk->dsc + NK_UN2);
}

pdx_status pdx_nk_jv_dev(pdx_ctx *c, const double *v, double *jv, double scale)
{
    pdx_nk *k = c->nk;
    const double sc = scale * std::cbrt(2.220446049250313e-16);
    PDX_TRY(multidot_dev(c, 1, v, v, k->dsc + NK_VN2));
    hipLaunchKernelGGL(k_perturb, grid1(c->n), dim3(256), 0, c->stream, c->n, c->Xr, c->u, v,
                       k->dsc + NK_UN2, k->dsc + NK_VN2, sc, k->recp, k->recm);
    PDX_HIP(c, hipGetLastError());
    PDX_TRY(force_scaled(c, k->recp, k->fp));
    PDX_TRY(force_scaled(c, k->recm, k->fm));
    hipLaunchKernelGGL(k_jv_combine, grid1(k->m), dim3(256), 0, c->stream, k->m, k->fp, k->fm,
                       k->dsc + NK_UN2, k->dsc + NK_VN2, sc, k->cmask, v, jv);
    PDX_HIP(c, hipGetLastError());
    k->n_jv++;
    return PDX_OK;
}

// ----------------------------------------------------------------------------------------
// Two-level preconditioner.  Aggregates are runs of `agg` consecutive Morton points
// (compact boxes).  Every point carries nf basis fields phi_i[f] (3-vectors, zero on
// constrained dofs); the coarse space of aggregate I holds those nf fields restricted to
// I, orthonormalised per aggregate.  The translations (nf = 3, phi_i = e_x, e_y, e_z) give
// the classical aggregation space; the learned preconditioner appends the fields of its
// network (pdx_gnn.hip).  The Galerkin operator A_c = P^T J P comes from the analytic
// linearisation of the bond-based residual (the bond stiffness of the block Jacobi) in
// two passes: Y_i,J = sum_{j in J} J_ij phi_j per point and neighbouring aggregate, then
// A_c[(I,k),(J,l)] = sum_{i in I} phi_i[k] . Y_i,J[l].  -A_c is symmetric positive
// definite: it is factored and inverted once per load step (rocSOLVER potrf + potri) and
// kept as float32 for the per-iteration product.
// z = D^-1 v + P A_c^-1 P^T v.
// ----------------------------------------------------------------------------------------
#define NK_TP 16    // aggregates a point's bonds may reach (slot 0: its own)
#define NK_FCH 4    // basis fields per pass of the Y kernel

// slots of a point: slot 0 is its own aggregate, the others in order of appearance along
// its (neighbour-sorted) bonds
__device__ __forceinline__ int nk_slot(int32_t *pj, int &ns, int64_t J, int *overflow)
{
    for (int t = 0; t < ns; t++)
        if (pj[t] == J) return t;
    if (ns >= NK_TP) {
        *overflow = 1;
        return -1;
    }
    pj[ns] = (int32_t)J;
    return ns++;
}

// Y[(i*TP + t)*nf + f][d] = sum over the bonds of i into slot t's aggregate of C_ij phi_j[f],
// plus -sum_j C_ij phi_i[f] in slot 0 (the diagonal block J_ii).  C_ij is the bond's
// stiffness dr_i/du_j = V_i bc V_j [(1/L) n n^T + (eps/l)(I - n n^T)].
__global__ __launch_bounds__(256) void k_coarse_y(int64_t n, int agg, int nf, const int64_t *slice_ptr,
                                                  const int32_t *cnt, const uint32_t *bw,
                                                  const double *L, const double4 *rec, double bc,
                                                  const pdx_pset *ps, const uint8_t *pid,
                                                  const double *phi, int32_t *pJ, double *Y,
                                                  int *overflow)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t I = i / agg;
    const int64_t b0 = slice_ptr[i >> 6] + (i & 63);
    const double4 ri = rec[i];
    int32_t pj[NK_TP];
    for (int f0 = 0; f0 < nf; f0 += NK_FCH) {
        const int fn = min(NK_FCH, nf - f0);
        int ns = 1;
        pj[0] = (int32_t)I;
        double S[3][NK_FCH], T[9];
        for (int q = 0; q < 9; q++) T[q] = 0.0;
        for (int d = 0; d < 3; d++)
            for (int q = 0; q < NK_FCH; q++) S[d][q] = 0.0;
        int64_t Jc = -1;
        auto flush = [&]() {
            if (Jc < 0) return;
            const int t = nk_slot(pj, ns, Jc, overflow);
            if (t >= 0)
                for (int q = 0; q < fn; q++)
                    for (int d = 0; d < 3; d++) {
                        double *y = Y + ((i * NK_TP + t) * nf + f0 + q) * 3 + d;
                        *y += S[d][q];
                    }
            for (int d = 0; d < 3; d++)
                for (int q = 0; q < NK_FCH; q++) S[d][q] = 0.0;
        };
        for (int t = 0; t < cnt[i]; t++) {
            const int64_t b = b0 + (int64_t)t * 64;
            const uint32_t wd = bw[b];
            if (!(wd & PDX_BW_ACT)) continue;
            const int64_t j = wd & PDX_BW_NBR;
            const int64_t J = j / agg;
            if (J != Jc) {
                flush();
                Jc = J;
            }
            const double4 rj = rec[j];
            const double dx = rj.x - ri.x, dy = rj.y - ri.y, dz = rj.z - ri.z;
            const double l = sqrt((dx * dx + dy * dy) + dz * dz);
            const double nv[3] = {dx / l, dy / l, dz / l};
            const double Lb = L[b];
            const double eps = (l - Lb) / Lb;
            const double bcb = ps ? (ps[pid[i]].bc + ps[pid[j]].bc) / 2.0 : bc;
            const double a = ri.w * bcb * rj.w / Lb, g = ri.w * bcb * rj.w * eps / l;
            double C[9];
            for (int d = 0; d < 3; d++)
                for (int e = 0; e < 3; e++) C[3 * d + e] = (a - g) * nv[d] * nv[e] + (d == e ? g : 0.0);
            for (int q = 0; q < 9; q++) T[q] -= C[q];
            const double *pjf = phi + (j * nf + f0) * 3;
            for (int q = 0; q < fn; q++)
                for (int d = 0; d < 3; d++)
                    S[d][q] += (C[3 * d] * pjf[3 * q] + C[3 * d + 1] * pjf[3 * q + 1]) + C[3 * d + 2] * pjf[3 * q + 2];
        }
        flush();
        // diagonal block into slot 0
        const double *pif = phi + (i * nf + f0) * 3;
        for (int q = 0; q < fn; q++)
            for (int d = 0; d < 3; d++) {
                double *y = Y + ((i * NK_TP) * nf + f0 + q) * 3 + d;
                *y += (T[3 * d] * pif[3 * q] + T[3 * d + 1] * pif[3 * q + 1]) + T[3 * d + 2] * pif[3 * q + 2];
            }
        if (f0 == 0)
            for (int t = 0; t < NK_TP; t++) pJ[i * NK_TP + t] = t < ns ? pj[t] : -1;
    }
}

// A_c[(I,k),(J,l)] = sum_{i in I} phi_i[k] . Y_i,J[l]: one workgroup per (aggregate,
// neighbour-aggregate slot), one thread per (k, l)
__global__ __launch_bounds__(256) void k_coarse_gather(int64_t n, int agg, int nf, int maxs,
                                                       const int32_t *aslot, const int32_t *pJ,
                                                       const double *phi, const double *Y,
                                                       int64_t mc, double *Ac)
{
    const int64_t I = blockIdx.x / maxs;
    const int s = (int)(blockIdx.x % maxs);
    const int32_t J = aslot[I * maxs + s];
    if (J < 0) return;
    const int kk = threadIdx.x / nf, ll = threadIdx.x % nf;
    const bool on = (int)threadIdx.x < nf * nf;
    const int64_t p0 = I * agg, p1 = min(n, p0 + agg);
    double acc = 0.0;
    for (int64_t i = p0; i < p1; i++) {
        int t = -1;
        for (int q = 0; q < NK_TP; q++)
            if (pJ[i * NK_TP + q] == J) {
                t = q;
                break;
            }
        if (t < 0 || !on) continue;
        const double *pk = phi + (i * nf + kk) * 3;
        const double *yl = Y + ((i * NK_TP + t) * nf + ll) * 3;
        acc += (pk[0] * yl[0] + pk[1] * yl[1]) + pk[2] * yl[2];
    }
    if (on) Ac[(I * nf + kk) * mc + (int64_t)J * nf + ll] = acc;
}

// the number of aggregates (its own included) that the bonds of a point reach, maximised
// over the points into *out
__global__ void k_coarse_reach(int64_t n, int agg, const int64_t *slice_ptr, const int32_t *cnt,
                               const uint32_t *bw, int *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t I = i / agg;
    const int64_t b0 = slice_ptr[i >> 6] + (i & 63);
    int runs = 0;
    bool own = false;
    int64_t Jc = -1;
    for (int t = 0; t < cnt[i]; t++) {
        const int64_t J = (int64_t)(bw[b0 + (int64_t)t * 64] & PDX_BW_NBR) / agg;
        if (J != Jc) {
            runs++;
            own |= (J == I);
            Jc = J;
        }
    }
    atomicMax(out, runs + (own ? 0 : 1));
}

// neighbouring aggregates: touch[I * nagg + J] = 1 if a point of I has a bond slot into J
__global__ void k_coarse_touch(int64_t n, int agg, int64_t nagg, const int32_t *pJ, uint8_t *touch)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t I = i / agg;
    for (int t = 0; t < NK_TP; t++) {
        const int32_t J = pJ[i * NK_TP + t];
        if (J >= 0) touch[I * nagg + J] = 1;
    }
}

// B = -A_c with an identity row for a coarse dof whose field vanishes on the aggregate
__global__ void k_coarse_negate(int64_t mc, double *A)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mc * mc) return;
    const int64_t r = e / mc, cc = e - r * mc;
    double v = -A[e];
    if (r == cc && v == 0.0) v = 1.0;
    A[e] = v;
}

// the inverse of -A_c (lower triangle, column-major, from potri) -> A_c^-1 in float32, both
// triangles
__global__ void k_coarse_finish(int64_t mc, const double *B, float *Ainv)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mc * mc) return;
    const int64_t r = e / mc, cc = e - r * mc;
    const double v = r >= cc ? B[cc * mc + r] : B[r * mc + cc];
    Ainv[e] = (float)(-v);
}

// basis fields of every point: the three translations, then `nl` learned fields (float,
// 3 per field), zero on constrained dofs
__global__ void k_coarse_basis(int64_t n, int nl, const float *learned, const uint8_t *cmask,
                               double *phi)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int nf = 3 + nl;
    double *p = phi + i * nf * 3;
    for (int f = 0; f < 3; f++)
        for (int d = 0; d < 3; d++) p[3 * f + d] = (d == f && !cmask[3 * i + d]) ? 1.0 : 0.0;
    for (int f = 0; f < nl; f++)
        for (int d = 0; d < 3; d++)
            p[3 * (3 + f) + d] = cmask[3 * i + d] ? 0.0 : (double)learned[i * 3 * nl + 3 * f + d];
}

// per aggregate: Gram-Schmidt (twice) of its nf fields over its points; a field whose
// remainder is below 1e-6 of its norm is dropped (zeroed).  One workgroup per aggregate.
#define NK_MAXF 16
__global__ __launch_bounds__(256) void k_coarse_orth(int64_t n, int agg, int nf, double *phi)
{
    __shared__ double red[256];
    const int64_t I = blockIdx.x;
    const int64_t p0 = I * agg, p1 = min(n, p0 + agg);
    for (int f = 0; f < nf; f++) {
        double nrm0 = 0.0;
        for (int pass = 0; pass < 3; pass++) {
            // pass 0: the field's norm; passes 1, 2: remove the earlier fields
            for (int g = 0; g < (pass == 0 ? 1 : f); g++) {
                double s = 0.0;
                for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
                    const double *a = phi + (p * nf + f) * 3;
                    const double *b = phi + (p * nf + (pass == 0 ? f : g)) * 3;
                    s += (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
                }
                red[threadIdx.x] = s;
                __syncthreads();
                for (int o = 128; o > 0; o >>= 1) {
                    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
                    __syncthreads();
                }
                const double dot = red[0];
                __syncthreads();
                if (pass == 0) {
                    nrm0 = sqrt(dot);
                } else {
                    for (int64_t p = p0 + threadIdx.x; p < p1; p += 256)
                        for (int d = 0; d < 3; d++)
                            phi[(p * nf + f) * 3 + d] -= dot * phi[(p * nf + g) * 3 + d];
                }
                __syncthreads();
            }
        }
        double s = 0.0;
        for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
            const double *a = phi + (p * nf + f) * 3;
            s += (a[0] * a[0] + a[1] * a[1]) + a[2] * a[2];
        }
        red[threadIdx.x] = s;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        const double nr = sqrt(red[0]);
        __syncthreads();
        const double sc = (nr > 1e-6 * nrm0 && nr > 0.0) ? 1.0 / nr : 0.0;
        for (int64_t p = p0 + threadIdx.x; p < p1; p += 256)
            for (int d = 0; d < 3; d++) phi[(p * nf + f) * 3 + d] *= sc;
        __syncthreads();
    }
}

// rc[I*nf + f] = sum_{i in I} phi_i[f] . v_i; one workgroup per aggregate
__global__ __launch_bounds__(256) void k_coarse_restrict(int64_t n, int agg, int nf, const double *phi,
                                                         const double *v, double *rc)
{
    __shared__ double sh[NK_MAXF][256];
    const int64_t I = blockIdx.x;
    const int64_t p0 = I * agg, p1 = min(n, p0 + agg);
    double s[NK_MAXF];
    for (int f = 0; f < NK_MAXF; f++) s[f] = 0.0;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
        const double v0 = v[3 * p], v1 = v[3 * p + 1], v2 = v[3 * p + 2];
        for (int f = 0; f < nf; f++) {
            const double *a = phi + (p * nf + f) * 3;
            s[f] += (a[0] * v0 + a[1] * v1) + a[2] * v2;
        }
    }
    for (int f = 0; f < nf; f++) sh[f][threadIdx.x] = s[f];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int f = 0; f < nf; f++) sh[f][threadIdx.x] += sh[f][threadIdx.x + o];
        __syncthreads();
    }
    if ((int)threadIdx.x < nf) rc[I * nf + threadIdx.x] = sh[threadIdx.x][0];
}

// zc = Ainv rc: one wavefront per row, coalesced along the row, shuffle reduction
__global__ __launch_bounds__(64) void k_coarse_matvec(int64_t mc, const float *Ainv, const double *rc,
                                                      double *zc)
{
    const int64_t r = blockIdx.x;
    const float *row = Ainv + r * mc;
    double s = 0.0;
    for (int64_t k = threadIdx.x; k < mc; k += 64) s += (double)row[k] * rc[k];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (threadIdx.x == 0) zc[r] = s;
}

// z = a D^-1 v + b P zc  (a = 0: the prolongation alone, added to z when b is given and
// `add` is set)
__global__ void k_apply_two_level(int64_t n, int agg, int nf, const double *dinv, const double *v,
                                  const double *zc, const double *phi, double alpha, int add,
                                  double *z)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double o[3] = {0.0, 0.0, 0.0};
    if (add) {
        o[0] = z[3 * i];
        o[1] = z[3 * i + 1];
        o[2] = z[3 * i + 2];
    } else {
        const double *R = dinv + 9 * i;
        const double a = v[3 * i], b = v[3 * i + 1], cc = v[3 * i + 2];
        o[0] = R[0] * a + R[3] * b + R[6] * cc;
        o[1] = R[1] * a + R[4] * b + R[7] * cc;
        o[2] = R[2] * a + R[5] * b + R[8] * cc;
    }
    const int64_t I = i / agg;
    for (int f = 0; f < nf; f++) {
        const double w = alpha * zc[I * nf + f];
        const double *p = phi + (i * nf + f) * 3;
        o[0] += w * p[0];
        o[1] += w * p[1];
        o[2] += w * p[2];
    }
    for (int d = 0; d < 3; d++) z[3 * i + d] = o[d];
}

// aggregate size: pdx_nk_agg_size(n), doubled until no point's bonds reach more than
// NK_TP aggregates (its own included)
pdx_status pdx_nk_aggregates(pdx_ctx *c)
{
    pdx_nk *k = c->nk;
    if (k->agg) return PDX_OK;
    const int64_t n = c->n;
    if (!k->dflag) PDX_TRY(pdx_alloc(c, &k->dflag, 4));
    for (int64_t agg = pdx_nk_agg_size(n);; agg *= 2) {
        PDX_HIP(c, hipMemsetAsync(k->dflag, 0, sizeof(int), c->stream));
        hipLaunchKernelGGL(k_coarse_reach, grid1(n), dim3(256), 0, c->stream, n, (int)agg, c->slice_ptr,
                           c->cnt, c->bw, k->dflag);
        PDX_HIP(c, hipGetLastError());
        int reach = 0;
        PDX_HIP(c, hipMemcpyAsync(&reach, k->dflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        PDX_HIP(c, hipStreamSynchronize(c->stream));
        if (reach <= NK_TP || agg >= n) {
            if (reach > NK_TP) return pdx_fail(c, PDX_ERR_UNSUPPORTED, "a point reaches more than %d aggregates", NK_TP);
            k->agg = (int)agg;
            k->nagg = (n + agg - 1) / agg;
            return PDX_OK;
        }
    }
}

static pdx_status nk_rocsolver(pdx_ctx *c, rocblas_status st, const char *what)
{
    if (st != rocblas_status_success)
        return pdx_fail(c, PDX_ERR_HIP, "%s failed (rocblas status %d)", what, (int)st);
    return PDX_OK;
}

// the coarse space, operator and inverse at the current state.  learned_fields: the
// network's fields (float, 3 * n_learned per point, internal order), or null for the
// translations alone.
static pdx_status build_coarse(pdx_ctx *c, const float *learned_fields, int n_learned)
{
    pdx_nk *k = c->nk;
    if (c->kind != PDX_BB) return pdx_fail(c, PDX_ERR_UNSUPPORTED, "the two-level preconditioner is built for BBMaterial");
    const int nf = 3 + (learned_fields ? n_learned : 0);
    if (nf > NK_MAXF) return pdx_fail(c, PDX_ERR_ARG, "at most %d coarse fields per aggregate", NK_MAXF);
    const int64_t n = c->n;
    hipStream_t s = c->stream;
    if (!k->pJ) {
        PDX_TRY(pdx_nk_aggregates(c));
        PDX_TRY(pdx_alloc(c, &k->pJ, (size_t)n * NK_TP));
        if (rocblas_create_handle(&k->rb) != rocblas_status_success)
            return pdx_fail(c, PDX_ERR_HIP, "rocblas_create_handle failed");
    }
    if (nf > k->nf_cap) {
        for (void *p : {(void *)k->phi, (void *)k->Ycs, (void *)k->Ac, (void *)k->Ainv, (void *)k->rc,
                        (void *)k->zc})
            if (p) (void)hipFree(p);
        k->phi = k->Ycs = k->Ac = k->rc = k->zc = nullptr;
        k->Ainv = nullptr;
        const int64_t mc = k->nagg * nf;
        PDX_TRY(pdx_alloc(c, &k->phi, (size_t)n * nf * 3));
        PDX_TRY(pdx_alloc(c, &k->Ycs, (size_t)n * NK_TP * nf * 3));
        PDX_TRY(pdx_alloc(c, &k->Ac, (size_t)(mc * mc)));
        PDX_TRY(pdx_alloc(c, &k->Ainv, (size_t)(mc * mc)));
        PDX_TRY(pdx_alloc(c, &k->rc, (size_t)mc));
        PDX_TRY(pdx_alloc(c, &k->zc, (size_t)mc));
        k->nf_cap = nf;
    }
    k->nf = nf;
    k->mc = k->nagg * nf;
    const int64_t mc = k->mc, mm = mc * mc;
    hipLaunchKernelGGL(k_coarse_basis, grid1(n), dim3(256), 0, s, n, nf - 3, learned_fields, k->cmask, k->phi);
    hipLaunchKernelGGL(k_coarse_orth, dim3((unsigned)k->nagg), dim3(256), 0, s, n, k->agg, nf, k->phi);
    PDX_HIP(c, hipMemsetAsync(k->Ycs, 0, sizeof(double) * n * NK_TP * nf * 3, s));
    PDX_HIP(c, hipMemsetAsync(k->dflag, 0, sizeof(int) * 4, s));
    hipLaunchKernelGGL(k_coarse_y, grid1(n), dim3(256), 0, s, n, k->agg, nf, c->slice_ptr, c->cnt, c->bw,
                       c->L, c->rec[c->cur], c->P.bc, c->multi() ? c->d_psets : nullptr, c->d_pid,
                       k->phi, k->pJ, k->Ycs, k->dflag);
    PDX_HIP(c, hipGetLastError());
    int ovf = 0;
    PDX_HIP(c, hipMemcpyAsync(&ovf, k->dflag, sizeof(int), hipMemcpyDeviceToHost, s));
    PDX_HIP(c, hipStreamSynchronize(s));
    if (ovf) return pdx_fail(c, PDX_ERR_UNSUPPORTED, "a point reaches more than %d aggregates", NK_TP - 1);
    if (k->maxs == 0) {
        // neighbouring aggregates of every aggregate (the bond set of a Newton-Krylov body
        // does not change)
        const int64_t na = k->nagg;
        uint8_t *touch = nullptr;
        PDX_HIP(c, hipMalloc((void **)&touch, na * na));
        PDX_HIP(c, hipMemsetAsync(touch, 0, na * na, s));
        hipLaunchKernelGGL(k_coarse_touch, grid1(n), dim3(256), 0, s, n, k->agg, na, k->pJ, touch);
        std::vector<uint8_t> h(na * na);
        PDX_HIP(c, hipMemcpyAsync(h.data(), touch, na * na, hipMemcpyDeviceToHost, s));
        PDX_HIP(c, hipStreamSynchronize(s));
        (void)hipFree(touch);
        int maxs = 1;
        for (int64_t I = 0; I < na; I++) {
            int cntI = 0;
            for (int64_t J = 0; J < na; J++) cntI += h[I * na + J];
            maxs = std::max(maxs, cntI);
        }
        std::vector<int32_t> slots(na * maxs, -1);
        for (int64_t I = 0; I < na; I++) {
            int q = 0;
            for (int64_t J = 0; J < na; J++)
                if (h[I * na + J]) slots[I * maxs + q++] = (int32_t)J;
        }
        PDX_TRY(pdx_alloc(c, &k->aslot, slots.size()));
        PDX_HIP(c, hipMemcpy(k->aslot, slots.data(), sizeof(int32_t) * slots.size(), hipMemcpyHostToDevice));
        k->maxs = maxs;
    }
    PDX_HIP(c, hipMemsetAsync(k->Ac, 0, sizeof(double) * mm, s));
    hipLaunchKernelGGL(k_coarse_gather, dim3((unsigned)(k->nagg * k->maxs)), dim3(256), 0, s, n, k->agg, nf,
                       k->maxs, k->aslot, k->pJ, k->phi, k->Ycs, mc, k->Ac);
    hipLaunchKernelGGL(k_coarse_negate, grid1(mm), dim3(256), 0, s, mc, k->Ac);
    PDX_HIP(c, hipGetLastError());
    rocblas_set_stream(k->rb, s);
    PDX_TRY(nk_rocsolver(c, rocsolver_dpotrf(k->rb, rocblas_fill_lower, (rocblas_int)mc, k->Ac,
                                              (rocblas_int)mc, k->dflag + 1), "rocsolver_dpotrf"));
    PDX_TRY(nk_rocsolver(c, rocsolver_dpotri(k->rb, rocblas_fill_lower, (rocblas_int)mc, k->Ac,
                                              (rocblas_int)mc, k->dflag + 2), "rocsolver_dpotri"));
    int info[2] = {0, 0};
    PDX_HIP(c, hipMemcpyAsync(info, k->dflag + 1, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
    PDX_HIP(c, hipStreamSynchronize(s));
    if (info[0] != 0 || info[1] != 0)
        return pdx_fail(c, PDX_ERR_HIP, "the coarse operator is not definite (potrf info %d, potri info %d)",
                        info[0], info[1]);
    hipLaunchKernelGGL(k_coarse_finish, grid1(mm), dim3(256), 0, s, mc, k->Ac, k->Ainv);
    PDX_HIP(c, hipGetLastError());
    k->coarse_ready = true;
    return PDX_OK;
}

// z = D^-1 v + P A_c^-1 P^T v
pdx_status pdx_nk_two_level(pdx_ctx *c, const double *v, double *z)
{
    pdx_nk *k = c->nk;
    if (!k->coarse_ready) return pdx_fail(c, PDX_ERR_STATE, "coarse space not built");
    hipLaunchKernelGGL(k_coarse_restrict, dim3((unsigned)k->nagg), dim3(256), 0, c->stream, c->n, k->agg,
                       k->nf, k->phi, v, k->rc);
    hipLaunchKernelGGL(k_coarse_matvec, dim3((unsigned)k->mc), dim3(64), 0, c->stream, k->mc, k->Ainv, k->rc, k->zc);
    hipLaunchKernelGGL(k_apply_two_level, grid1(c->n), dim3(256), 0, c->stream, c->n, k->agg, k->nf, k->dinv, v,
                       k->zc, k->phi, 1.0, 0, z);
    PDX_HIP(c, hipGetLastError());
    return PDX_OK;
}

// z += alpha P A_c^-1 P^T v
pdx_status pdx_nk_coarse_add(pdx_ctx *c, const double *v, double *z, double alpha)
{
    pdx_nk *k = c->nk;
    if (!k->coarse_ready) return pdx_fail(c, PDX_ERR_STATE, "coarse space not built");
    hipLaunchKernelGGL(k_coarse_restrict, dim3((unsigned)k->nagg), dim3(256), 0, c->stream, c->n, k->agg,
                       k->nf, k->phi, v, k->rc);
    hipLaunchKernelGGL(k_coarse_matvec, dim3((unsigned)k->mc), dim3(64), 0, c->stream, k->mc, k->Ainv, k->rc, k->zc);
    hipLaunchKernelGGL(k_apply_two_level, grid1(c->n), dim3(256), 0, c->stream, c->n, k->agg, k->nf, k->dinv, v,
                       k->zc, k->phi, alpha, 1, z);
    PDX_HIP(c, hipGetLastError());
    return PDX_OK;
}

// z = M^-1 v for the chosen preconditioner
static pdx_status precond_apply(pdx_ctx *c, int which, const double *v, double *z)
{
    pdx_nk *k = c->nk;
    switch (which) {
    case PDX_PRE_JACOBI:
        hipLaunchKernelGGL(k_apply_block, grid1(c->n), dim3(256), 0, c->stream, c->n, k->dinv, v, z);
        PDX_HIP(c, hipGetLastError());
        return PDX_OK;
    case PDX_PRE_TWO_LEVEL: return pdx_nk_two_level(c, v, z);
    case PDX_PRE_LEARNED: return pdx_gnn_apply(c, v, z);  // two-level with the learned fields
    default:
        PDX_HIP(c, hipMemcpyAsync(z, v, sizeof(double) * k->m, hipMemcpyDeviceToDevice, c->stream));
        return PDX_OK;
    }
}

static pdx_status build_jacobi(pdx_ctx *c)
{
    if (c->kind != PDX_BB) return pdx_fail(c, PDX_ERR_UNSUPPORTED, "block Jacobi is built for BBMaterial");
    hipLaunchKernelGGL(k_bb_block_jacobi, grid1(c->n), dim3(256), 0, c->stream, c->n, c->slice_ptr,
                       c->cnt, c->bw, c->L, c->rec[c->cur], c->P.bc,
                       c->multi() ? c->d_psets : nullptr, c->d_pid, c->nk->cmask, c->nk->dinv);
    PDX_HIP(c, hipGetLastError());
    return PDX_OK;
}

pdx_status pdx_nk_build_jacobi(pdx_ctx *c)
{
    if (!c->nk) return pdx_fail(c, PDX_ERR_STATE, "Newton-Krylov workspace missing");
    return build_jacobi(c);
}

pdx_status pdx_nk_build_coarse(pdx_ctx *c)
{
    if (!c->nk) return pdx_fail(c, PDX_ERR_STATE, "Newton-Krylov workspace missing");
    return build_coarse(c, nullptr, 0);
}

pdx_status pdx_nk_build_coarse_fields(pdx_ctx *c, const float *fields, int n_fields)
{
    if (!c->nk) return pdx_fail(c, PDX_ERR_STATE, "Newton-Krylov workspace missing");
    return build_coarse(c, fields, n_fields);
}

// Flexible GMRES(restart) for J du = b, du in k->du (starts at zero).  Stops at
// ||res|| <= max(reltol ||b||, abstol) or after maxiter products.  Classical Gram-Schmidt
// with one reorthogonalisation pass; the coefficients and norms stay in device memory and
// the host reads them once per iteration for the Givens rotations.
static pdx_status fgmres(pdx_ctx *c, const double *b, int restart, int maxiter, double reltol,
                         double abstol, int precond, double scale, int *iters_out,
                         double *res_out)
{
    pdx_nk *k = c->nk;
    const int64_t m = k->m;
    const int R1 = PDX_MAX_RESTART + 1;
    double *hd = k->h_dev, *h2d = k->h_dev + R1, *hnd = k->h_dev + 2 * R1;  // device
    PDX_HIP(c, hipMemsetAsync(k->du, 0, sizeof(double) * m, c->stream));
    PDX_TRY(nk_refresh_unorm(c));
    double bnorm;
    PDX_TRY(norm2(c, b, &bnorm));
    const double thr = std::max(reltol * bnorm, abstol);
    int it = 0;
    PDX_HIP(c, hipMemcpyAsync(k->r, b, sizeof(double) * m, hipMemcpyDeviceToDevice, c->stream));
    double beta = bnorm;
    std::vector<double> H((size_t)(restart + 1) * restart), cs(restart), sn(restart), g(restart + 1),
        y(restart), hh(2 * R1 + 1);
    while (it < maxiter && beta > thr) {
        std::fill(H.begin(), H.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        hipLaunchKernelGGL(k_scale, grid1(m), dim3(256), 0, c->stream, m, 1.0 / beta, k->r, k->Vb);
        PDX_HIP(c, hipGetLastError());
        int j = 0;
        bool done = false;
        for (j = 0; j < restart; j++) {
            double *Vj = k->Vb + (int64_t)j * m;
            double *Zj = k->Zb + (int64_t)j * m;
            double *w = k->Vb + (int64_t)(j + 1) * m;
            PDX_TRY(precond_apply(c, precond, Vj, Zj));
            PDX_TRY(pdx_nk_jv_dev(c, Zj, w, scale));
            it++;
            if (pdx_gnn_collecting(c)) PDX_TRY(pdx_gnn_record_krylov(c, Zj, w));
            // classical Gram-Schmidt, twice, then the norm -- all on the device
            PDX_TRY(multidot_dev(c, j + 1, k->Vb, w, hd));
            hipLaunchKernelGGL(k_multiaxpy, grid1(m), dim3(256), 0, c->stream, m, j + 1, k->Vb, m, hd, w);
            PDX_TRY(multidot_dev(c, j + 1, k->Vb, w, h2d));
            hipLaunchKernelGGL(k_multiaxpy, grid1(m), dim3(256), 0, c->stream, m, j + 1, k->Vb, m, h2d, w);
            PDX_TRY(multidot_dev(c, 1, w, w, hnd));
            hipLaunchKernelGGL(k_scale_inv_sqrt, grid1(m), dim3(256), 0, c->stream, m, hnd, w, w);
            PDX_HIP(c, hipGetLastError());
            PDX_HIP(c, hipMemcpyAsync(hh.data(), k->h_dev, sizeof(double) * (2 * R1 + 1),
                                      hipMemcpyDeviceToHost, c->stream));
            PDX_HIP(c, hipStreamSynchronize(c->stream));
            for (int i = 0; i <= j; i++) H[i + (size_t)(restart + 1) * j] = hh[i] + hh[R1 + i];
            const double hn = std::sqrt(std::max(hh[2 * R1], 0.0));
            H[j + 1 + (size_t)(restart + 1) * j] = hn;
            // Givens
            double *Hc = &H[(size_t)(restart + 1) * j];
            for (int i = 0; i < j; i++) {
                const double t = cs[i] * Hc[i] + sn[i] * Hc[i + 1];
                Hc[i + 1] = -sn[i] * Hc[i] + cs[i] * Hc[i + 1];
                Hc[i] = t;
            }
            const double den = std::hypot(Hc[j], Hc[j + 1]);
            cs[j] = den != 0 ? Hc[j] / den : 1.0;
            sn[j] = den != 0 ? Hc[j + 1] / den : 0.0;
            Hc[j] = den;
            Hc[j + 1] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            if (std::fabs(g[j + 1]) <= thr || it >= maxiter || hn == 0) {
                j++;
                done = true;
                break;
            }
        }
        if (!done) j = restart;
        // y = H^-1 g (upper triangular j x j)
        for (int i = j - 1; i >= 0; i--) {
            double sum = g[i];
            for (int l = i + 1; l < j; l++) sum -= H[i + (size_t)(restart + 1) * l] * y[l];
            y[i] = sum / H[i + (size_t)(restart + 1) * i];
        }
        PDX_HIP(c, hipMemcpyAsync(k->h_dev, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_lincomb_add, grid1(m), dim3(256), 0, c->stream, m, j, k->Zb, m, k->h_dev, k->du);
        PDX_HIP(c, hipGetLastError());
        const double est = std::fabs(g[j]);
        if (est <= thr || it >= maxiter) {
            beta = est;
            break;
        }
        // true residual for the restart: r = b - J du
        PDX_TRY(pdx_nk_jv_dev(c, k->du, k->r, scale));
        it++;
        hipLaunchKernelGGL(k_sub_from, grid1(m), dim3(256), 0, c->stream, m, b, k->r);
        PDX_HIP(c, hipGetLastError());
        PDX_TRY(norm2(c, k->r, &beta));
    }
    *iters_out = it;
    *res_out = bnorm > 0 ? beta / bnorm : 0.0;
    return PDX_OK;
}

// ----------------------------------------------------------------------------------------
// C ABI
// ----------------------------------------------------------------------------------------
extern "C" pdx_status pdx_set_displacement_bcs(pdx_ctx *c, int64_t n, const int32_t *points,
                                               const int32_t *dims, const double *values)
{
    if (!c) return PDX_ERR_ARG;
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    if (n < 0 || (n > 0 && (!points || !dims || !values))) return pdx_fail(c, PDX_ERR_ARG, "invalid BC table");
    PDX_HIP(c, hipSetDevice(c->device));
    PDX_TRY(nk_ensure(c, 1));
    pdx_nk *k = c->nk;
    std::vector<int64_t> dof(n);
    std::vector<uint8_t> mask(3 * c->n, 0);
    for (int64_t q = 0; q < n; q++) {
        if (points[q] < 0 || points[q] >= c->n || dims[q] < 0 || dims[q] > 2)
            return pdx_fail(c, PDX_ERR_ARG, "BC point or dimension out of range");
        dof[q] = 3 * (int64_t)c->iperm[points[q]] + dims[q];
        if (mask[dof[q]]) return pdx_fail(c, PDX_ERR_ARG, "dof constrained twice");
        mask[dof[q]] = 1;
    }
    if (k->cdof) (void)hipFree(k->cdof);
    if (k->cval) (void)hipFree(k->cval);
    k->cdof = nullptr;
    k->cval = nullptr;
    PDX_TRY(pdx_alloc(c, &k->cdof, std::max<int64_t>(n, 1)));
    PDX_TRY(pdx_alloc(c, &k->cval, std::max<int64_t>(n, 1)));
    if (n > 0) {
        PDX_HIP(c, hipMemcpy(k->cdof, dof.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice));
        PDX_HIP(c, hipMemcpy(k->cval, values, sizeof(double) * n, hipMemcpyHostToDevice));
    }
    PDX_HIP(c, hipMemcpy(k->cmask, mask.data(), 3 * c->n, hipMemcpyHostToDevice));
    k->nc = n;
    return PDX_OK;
}

extern "C" pdx_status pdx_nk_coarse_info(pdx_ctx *c, int32_t *agg, int64_t *n_aggregates,
                                         int32_t *fields)
{
    if (!c) return PDX_ERR_ARG;
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    PDX_HIP(c, hipSetDevice(c->device));
    PDX_TRY(nk_ensure(c, 1));
    PDX_TRY(pdx_nk_aggregates(c));
    if (agg) *agg = c->nk->agg;
    if (n_aggregates) *n_aggregates = c->nk->nagg;
    if (fields) *fields = c->nk->coarse_ready ? c->nk->nf : 0;
    return PDX_OK;
}

extern "C" pdx_status pdx_residual(pdx_ctx *c, double *r_host, double *norm_out)
{
    if (!c) return PDX_ERR_ARG;
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    PDX_HIP(c, hipSetDevice(c->device));
    PDX_TRY(nk_ensure(c, 1));
    pdx_nk *k = c->nk;
    PDX_TRY(pdx_reset_flags(c));
    pdx_step_args a;
    PDX_TRY(pdx_fill_step_args(c, &a, 0, PDX_F_FORCE_ONLY | PDX_F_FP64, 0.0, nullptr));
    PDX_TRY(pdx_launch_step(c, a));
    hipLaunchKernelGGL(k_residual, grid1(k->m), dim3(256), 0, c->stream, c->n, c->bint, c->bext,
                       c->Xr, k->cmask, k->r);
    PDX_HIP(c, hipGetLastError());
    double nr;
    PDX_TRY(norm2(c, k->r, &nr));
    if (norm_out) *norm_out = nr;
    if (r_host) PDX_TRY(pdx_download_vec(c, k->r, r_host, 3, 3));
    return PDX_OK;
}

extern "C" pdx_status pdx_jv(pdx_ctx *c, const double *v_host, double *jv_host, double scale)
{
    if (!c || !v_host || !jv_host) return PDX_ERR_ARG;
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    PDX_HIP(c, hipSetDevice(c->device));
    PDX_TRY(nk_ensure(c, 1));
    pdx_nk *k = c->nk;
    PDX_TRY(pdx_upload_vec(c, v_host, k->t0, 3));
    PDX_TRY(nk_refresh_unorm(c));
    PDX_TRY(pdx_nk_jv_dev(c, k->t0, k->w, scale));
    return pdx_download_vec(c, k->w, jv_host, 3, 3);
}

extern "C" pdx_status pdx_nk_step(pdx_ctx *c, int64_t n, int64_t n_steps, const pdx_nk_opts *o,
                                  pdx_nk_stats *st)
{
    if (!c || !o) return PDX_ERR_ARG;
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    if (n_steps <= 0 || n < 1 || n > n_steps) return pdx_fail(c, PDX_ERR_ARG, "invalid load step");
    if (o->maxiter <= 0 || o->gmres_maxiter <= 0 || o->gmres_restart <= 0 ||
        o->gmres_restart > PDX_MAX_RESTART)
        return pdx_fail(c, PDX_ERR_ARG, "invalid Newton-Krylov options");
    if (o->precond < PDX_PRE_NONE || o->precond > PDX_PRE_TWO_LEVEL)
        return pdx_fail(c, PDX_ERR_ARG, "unknown preconditioner");
    PDX_HIP(c, hipSetDevice(c->device));
    PDX_TRY(nk_ensure(c, o->gmres_restart));
    pdx_nk *k = c->nk;
    const auto t0 = std::chrono::steady_clock::now();
    pdx_nk_stats S{};
    k->n_jv = 0;
    const double beta = (double)n / (double)n_steps;
    if (k->nc > 0) {
        hipLaunchKernelGGL(k_nk_bcs, grid1(k->nc), dim3(256), 0, c->stream, k->nc, k->cdof, k->cval,
                           beta, c->u, c->rec[c->cur], c->Xr);
        PDX_HIP(c, hipGetLastError());
    }
    const double scale = o->perturbation_scale > 0 ? o->perturbation_scale : 1.0;
    double r_first = -1.0;
    bool converged = false;
    for (int iter = 1; iter <= o->maxiter; iter++) {
        double rn;
        PDX_TRY(pdx_residual(c, nullptr, &rn));
        if (r_first < 0) r_first = rn;
        S.residual_final = rn;
        S.newton_iters = iter;
        const double thr = o->rel_tol > 0 ? o->rel_tol * r_first : o->tol;
        if (rn < thr) {
            converged = true;
            break;
        }
        if (iter == o->maxiter) break;
        // J du = -r
        hipLaunchKernelGGL(k_scale, grid1(k->m), dim3(256), 0, c->stream, k->m, -1.0, k->r, k->t0);
        PDX_HIP(c, hipGetLastError());
        const int pre = o->precond;
        if (pre != PDX_PRE_NONE) PDX_TRY(build_jacobi(c));
        // the coarse operator of the two-level and learned preconditioners: once per load
        // step (the learned one spans the translations and the network's fields)
        if (pre == PDX_PRE_TWO_LEVEL && iter == 1) PDX_TRY(build_coarse(c, nullptr, 0));
        if (pre == PDX_PRE_LEARNED && iter == 1) {
            PDX_TRY(pdx_gnn_prepare(c));
            const float *fl = nullptr;
            int nl = 0;
            if (pdx_gnn_active(c)) PDX_TRY(pdx_gnn_fields(c, &fl, &nl));
            PDX_TRY(build_coarse(c, fl, nl));
        }
        int its = 0;
        double rel = 0.0;
        if (pre == PDX_PRE_LEARNED) {
            // with the network switched off the learned preconditioner is the two-level one
            // of the translations: nothing to fall back to
            const bool net = pdx_gnn_active(c);
            PDX_TRY(fgmres(c, k->t0, o->gmres_restart, o->gmres_maxiter, o->gmres_reltol,
                           o->gmres_abstol, PDX_PRE_LEARNED, scale, &its, &rel));
            S.krylov_iters += its;
            // fallback: a learned solve that did not reach the tolerance and did not reduce
            // the linear residual tenfold is redone with the two-level preconditioner of the
            // translations, and the network is not applied again in this context
            const double tol_lin = std::max(o->gmres_reltol, o->gmres_abstol / std::max(rn, 1e-300));
            if (net && !(rel <= tol_lin * 1.0000001) && !(rel <= 0.1)) {
                S.precond_fallbacks++;
                pdx_gnn_disable(c);
                PDX_TRY(build_coarse(c, nullptr, 0));
                PDX_TRY(fgmres(c, k->t0, o->gmres_restart, o->gmres_maxiter, o->gmres_reltol,
                               o->gmres_abstol, PDX_PRE_TWO_LEVEL, scale, &its, &rel));
                S.krylov_iters += its;
            }
        } else {
            PDX_TRY(fgmres(c, k->t0, o->gmres_restart, o->gmres_maxiter, o->gmres_reltol,
                           o->gmres_abstol, pre, scale, &its, &rel));
            S.krylov_iters += its;
        }
        if (k->verbose)
            fprintf(stderr, "[pdx nk] step %lld newton %d: |r| %.6e, krylov %d, linear rel %.3e\n",
                    (long long)n, iter, rn, its, rel);
        if (pdx_gnn_collecting(c)) PDX_TRY(pdx_gnn_record(c, k->t0, k->du));
        hipLaunchKernelGGL(k_nk_update, grid1(k->m), dim3(256), 0, c->stream, c->n, k->du, k->cmask,
                           c->u, c->rec[c->cur]);
        PDX_HIP(c, hipGetLastError());
    }
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    S.residual_first = r_first;
    S.jv_evals = (int32_t)k->n_jv;
    S.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (st) *st = S;
    if (!converged)
        return pdx_fail(c, PDX_ERR_NOT_CONVERGED,
                        "Newton-Krylov solver did not converge after max iterations!");
    return PDX_OK;
}
