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
#include <cstdlib>

#include "pdx_internal.h"
#include "pdx_nk.h"

#define RED_BLOCKS 512
#define CGS_MAX_BLOCKS 2048  // partial sums of the Gram-Schmidt kernels: at most this many blocks
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
                    (void *)k->phis, (void *)k->Ycs, (void *)k->pJ, (void *)k->aslot, (void *)k->dflag,
                    (void *)k->vparts, (void *)k->zparts})
        if (p) (void)hipFree(p);
    if (k->rb) rocblas_destroy_handle(k->rb);
    if (k->gm) (void)hipFree(k->gm);
    if (k->flag_host) (void)hipHostFree(k->flag_host);
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
    // columns queued past convergence carry y = 0 and take no part, whatever they hold
    for (int k = 0; k < nv; k++)
        if (y[k] != 0.0) s += y[k] * Z[k * ldz + i];
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
// positions X + (u + s v) and X + (u - s v) with the volume in .w
// (store_state_and_apply_perturbation!, store_pos_force_and_apply_neg_perturbation!)
__global__ void k_perturb(int64_t n, const double4 *Xr, const double *u, const double *v,
                          const double *un2, const double *vn2, double sc, double4 *rp, double4 *rm)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = pdx_nk_perturbation(un2, vn2, sc);
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
    const double s = pdx_nk_perturbation(un2, vn2, sc);
    jv[t] = cmask[t] ? v[t] : (fp[t] - fm[t]) / (2.0 * s);
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

// Gram-Schmidt passes over a 3n vector w and the first vectors of V (leading dimension ldv),
// partial-sum kernels (one row of partial sums per block) that k_reduce_partials or the
// Arnoldi step reduce:
//   k_cgs_dots:  partial = V_k . w                        (k < nv)
//   k_cgs_axdn:  w -= V h;  partial = V_k . w, then w . w  (k < nv; the V values of a dof
//                stay in registers between the update and the dots when nv <= 32)
// (the next basis vector (w - V h) / hn is formed by k_next_basis)
// The 32 running sums of a lane are reduced over the wave by pdx_wave_reduce32 (a halving
// butterfly); all indices into the register arrays are compile-time constants, so the
// arrays stay in registers.
#define CGS_REG 32
// threads per block of the Gram-Schmidt kernels (one dof per thread).  256 keeps five blocks
// per CU resident (78-95 VGPRs), so a 300 k-dof vector's 1172 blocks run in one round; 512
// and 1024 leave 4x fewer partial sums to reduce but measured 0.116 and 0.125 ms per
// block-Jacobi Krylov iteration against 0.114 ms.
#ifndef CGS_T
#define CGS_T 256
#endif
#define CGS_NW (CGS_T / 64)
// one dof per thread (grid-stride beyond CGS_MAX_BLOCKS blocks): the loads of all vectors
// of a dof are in flight at once
static inline int cgs_blocks(int64_t m) { return (int)std::min<int64_t>((m + CGS_T - 1) / CGS_T, CGS_MAX_BLOCKS); }

// the block's sums 0..kn-1 (after pdx_wave_reduce32) into column q of the partial sums
// (partial[q * nblk + block]: the reductions read a sum's partials contiguously), in a fixed
// order
__device__ __forceinline__ void cgs_block_store(const double (&s)[CGS_REG], int kn,
                                                double (*sh)[CGS_REG], double *partial, int q0)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if ((lane & 1) == 0) sh[wv][lane >> 1] = s[0];
    __syncthreads();
    if ((int)threadIdx.x < kn) {
        const int q = threadIdx.x;
        double t = 0.0;
#pragma unroll
        for (int b = 0; b < CGS_NW; b++) t += sh[b][q];
        partial[(int64_t)(q0 + q) * gridDim.x + blockIdx.x] = t;
    }
    __syncthreads();
}

__device__ __forceinline__ double cgs_wave_sum(double t)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
    return t;
}

// the block's partial of one more sum (x of every thread) into column q
__device__ __forceinline__ void cgs_block_store1(double x, double *partial, int q)
{
    __shared__ double sh1[CGS_NW];
    x = cgs_wave_sum(x);
    if ((threadIdx.x & 63) == 0) sh1[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (int b = 0; b < CGS_NW; b++) t += sh1[b];
        partial[(int64_t)q * gridDim.x + blockIdx.x] = t;
    }
}

// partial (column q, one per block) = V_q . w for q < nv
__global__ __launch_bounds__(CGS_T) void k_cgs_dots(int64_t m, int nv, const double *V, int64_t ldv,
                                                  const double *w, double *partial)
{
    __shared__ double sh[CGS_NW][CGS_REG];
    const int64_t i0 = (int64_t)blockIdx.x * CGS_T, stride = (int64_t)gridDim.x * CGS_T;
    if (m <= stride && nv <= CGS_REG) {
        // one dof per thread: the products go straight into the register array
        const int64_t i = i0 + threadIdx.x;
        double s[CGS_REG];
        const double wi = i < m ? w[i] : 0.0;
#pragma unroll
        for (int q = 0; q < CGS_REG; q++) s[q] = (q < nv && i < m) ? V[(int64_t)q * ldv + i] * wi : 0.0;
        pdx_wave_reduce32(s);
        cgs_block_store(s, nv, sh, partial, 0);
        return;
    }
    for (int k0 = 0; k0 < nv; k0 += CGS_REG) {
        const int kn = min(CGS_REG, nv - k0);
        const double *Vk = V + (int64_t)k0 * ldv;
        double s[CGS_REG];
#pragma unroll
        for (int q = 0; q < CGS_REG; q++) s[q] = 0.0;
        for (int64_t i = i0 + threadIdx.x; i < m; i += stride) {
            const double wi = w[i];
#pragma unroll
            for (int q = 0; q < CGS_REG; q++)
                if (q < kn) s[q] += Vk[(int64_t)q * ldv + i] * wi;
        }
        pdx_wave_reduce32(s);
        cgs_block_store(s, kn, sh, partial, k0);
    }
}

// w -= V h; partial columns q < nv: V_q . w (the updated w), column nv: w . w.  One dof
// per thread (m <= CGS_T x blocks) and nv <= 32: the products overwrite the loaded V values,
// so only one 32-wide register array is live; h is read from LDS.
__global__ __launch_bounds__(CGS_T) void k_cgs_axdn(int64_t m, int nv, const double *h, const double *V,
                                                  int64_t ldv, double *w, double *partial)
{
    __shared__ double sh[CGS_NW][CGS_REG];
    __shared__ double hs[CGS_REG];
    if (threadIdx.x < CGS_REG) hs[threadIdx.x] = (int)threadIdx.x < nv ? h[threadIdx.x] : 0.0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * CGS_T + threadIdx.x;
    double v[CGS_REG];
    double x = 0.0;
#pragma unroll
    for (int q = 0; q < CGS_REG; q++) v[q] = 0.0;
    if (i < m) {
#pragma unroll
        for (int q = 0; q < CGS_REG; q++)
            if (q < nv) v[q] = V[(int64_t)q * ldv + i];
        x = w[i];
#pragma unroll
        for (int q = 0; q < CGS_REG; q++)
            if (q < nv) x -= hs[q] * v[q];
        w[i] = x;
    }
#pragma unroll
    for (int q = 0; q < CGS_REG; q++) v[q] *= x;
    pdx_wave_reduce32(v);
    cgs_block_store(v, nv, sh, partial, 0);
    cgs_block_store1(x * x, partial, nv);
}

// the same for several dofs per thread or a long basis: the update first, then the dots in
// chunks (V read again)
__global__ __launch_bounds__(CGS_T) void k_cgs_axdn_loop(int64_t m, int nv, const double *h, const double *V,
                                                       int64_t ldv, double *w, double *partial)
{
    __shared__ double sh[CGS_NW][CGS_REG];
    const int64_t i0 = (int64_t)blockIdx.x * CGS_T, stride = (int64_t)gridDim.x * CGS_T;
    double nrm = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < m; i += stride) {
        double x = w[i];
        for (int k = 0; k < nv; k++) x -= h[k] * V[(int64_t)k * ldv + i];
        w[i] = x;
        nrm += x * x;
    }
    for (int k0 = 0; k0 < nv; k0 += CGS_REG) {
        const int kn = min(CGS_REG, nv - k0);
        const double *Vk = V + (int64_t)k0 * ldv;
        double s[CGS_REG];
#pragma unroll
        for (int q = 0; q < CGS_REG; q++) s[q] = 0.0;
        for (int64_t i = i0 + threadIdx.x; i < m; i += stride) {
            const double wi = w[i];
#pragma unroll
            for (int q = 0; q < CGS_REG; q++)
                if (q < kn) s[q] += Vk[(int64_t)q * ldv + i] * wi;
        }
        pdx_wave_reduce32(s);
        cgs_block_store(s, kn, sh, partial, k0);
    }
    cgs_block_store1(nrm, partial, nv);
}

// out[q] = sum of the nblk partials of column q (one workgroup per q)
__global__ __launch_bounds__(256) void k_reduce_cols(int nblk, const double *partial, double *out)
{
    __shared__ double sh[4];
    const double *col = partial + (int64_t)blockIdx.x * nblk;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) s += col[b];
    s = cgs_wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}



// doubles of the small GMRES state for restart length R (H, cs, sn, g, y, h1, h2, hn and
// two ints)
static size_t nk_gm_doubles(int R) { return (size_t)(R + 1) * R + 2 * R + (R + 1) + R + 2 * (R + 1) + 2 + 2; }

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
        PDX_TRY(pdx_alloc(c, &k->red, (size_t)std::max(RED_BLOCKS, CGS_MAX_BLOCKS) * (PDX_MAX_RESTART + 2)));
        PDX_TRY(pdx_alloc(c, &k->vparts, CGS_MAX_BLOCKS));
        PDX_TRY(pdx_alloc(c, &k->zparts, (size_t)((n + 63) / 64)));
        PDX_TRY(pdx_alloc(c, &k->h_dev, 2 * (PDX_MAX_RESTART + 1) + 8));
        PDX_TRY(pdx_alloc(c, &k->dsc, 8));
        PDX_HIP(c, hipMemset(k->dsc, 0, sizeof(double) * 8));
        PDX_TRY(pdx_alloc(c, &k->dinv, 9 * n));
        PDX_TRY(pdx_alloc(c, &k->gm, nk_gm_doubles(PDX_MAX_RESTART)));
        // J v in one kernel where the layout allows; PDX_JV_TWO_KERNELS=1 keeps the two force
        // evaluations of the step kernel (the same numbers, for comparison)
        const char *two = std::getenv("PDX_JV_TWO_KERNELS");
        k->jv_fused = pdx_bb_jv_fusable(c) && !(two && two[0] == '1');
        PDX_HIP(c, hipHostMalloc((void **)&k->flag_host, sizeof(int) * PDX_NK_SLOTS, hipHostMallocCoherent));
        PDX_HIP(c, hipHostGetDevicePointer((void **)&k->flag_dev, k->flag_host, 0));
        std::memset(k->flag_host, 0, sizeof(int) * PDX_NK_SLOTS);
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
static pdx_status nk_norm2_dev(pdx_ctx *c, const double *x, double *out_dev);
static pdx_status nk_refresh_unorm(pdx_ctx *c) { return nk_norm2_dev(c, c->u, c->nk->dsc + NK_UN2); }

// sum of np partial sums into out[0] (the order of pdx_block_sum_parts at 1024 threads, as
// the one-kernel J v forms ||v||^2)
__global__ __launch_bounds__(1024) void k_sum_parts(const double *parts, int np, double *out)
{
    __shared__ double sh[16];
    const double t = pdx_block_sum_parts(parts, np, sh);
    if (threadIdx.x == 0) out[0] = t;
}

// J v; ||v||^2 is the sum of the np partial sums `parts` when given (the preconditioner's
// kernels leave them for z), else formed here
pdx_status pdx_nk_jv_dev(pdx_ctx *c, const double *v, double *jv, double scale, const double *parts,
                         int np)
{
    pdx_nk *k = c->nk;
    const double sc = scale * std::cbrt(2.220446049250313e-16);
    if (!parts) {
        const int nb = cgs_blocks(k->m);
        hipLaunchKernelGGL(k_cgs_dots, dim3(nb), dim3(CGS_T), 0, c->stream, k->m, 1, v, k->m, v, k->vparts);
        parts = k->vparts;
        np = nb;
    }
    if (k->jv_fused) {
        PDX_TRY(pdx_launch_bb_jv(c, c->u, v, k->dsc + NK_UN2, parts, np, sc, k->cmask, jv));
        k->n_jv++;
        return PDX_OK;
    }
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(1024), 0, c->stream, parts, np, k->dsc + NK_VN2);
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

// phis[((c nf + f) 3 + d) 64 + l] = component d of field f of point 64 c + l (zero past n): the
// basis by 64-point chunk, field, component and lane, the layout in which the per-iteration
// kernels read it with consecutive lanes on consecutive addresses
__global__ __launch_bounds__(256) void k_coarse_soa(int64_t n, int nf, const double *phi, double *phis)
{
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (n + 63) / 64 * 64) return;
    const int64_t cc = p >> 6;
    const int l = (int)(p & 63);
    for (int f = 0; f < nf; f++)
        for (int d = 0; d < 3; d++)
            phis[((cc * nf + f) * 3 + d) * 64 + l] = p < n ? phi[(p * nf + f) * 3 + d] : 0.0;
}

// rc[I*nf + f] = sum_{i in I} phi_i[f] . v_i; one workgroup per aggregate (a whole number of
// 64-point chunks: the aggregate size is a power of two >= 64) with a wavefront per chunk (at
// most 16, which then take every 16th chunk) and a lane per point, then a lane-shuffle and a
// sum over the wavefronts in a fixed order
__global__ __launch_bounds__(1024) void k_coarse_restrict(int64_t n, int agg, int nf, const double *phis,
                                                          const double *v, double *rc)
{
    __shared__ double sh[16][NK_MAXF];
    const int64_t I = blockIdx.x;
    const int nw = (int)(blockDim.x >> 6);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int64_t nch = (n + 63) / 64;
    const int64_t c0 = I * (agg / 64), c1 = min(nch, (I + 1) * (agg / 64));
    double s[NK_MAXF];
#pragma unroll
    for (int f = 0; f < NK_MAXF; f++) s[f] = 0.0;
    for (int64_t cc = c0 + w; cc < c1; cc += nw) {
        const int64_t p = cc * 64 + l;
        double v0 = 0.0, v1 = 0.0, v2 = 0.0;
        if (p < n) {
            v0 = v[3 * p];
            v1 = v[3 * p + 1];
            v2 = v[3 * p + 2];
        }
        const double *b = phis + cc * nf * 3 * 64 + l;
#pragma unroll
        for (int f = 0; f < NK_MAXF; f++)
            if (f < nf) s[f] += (b[(3 * f) * 64] * v0 + b[(3 * f + 1) * 64] * v1) + b[(3 * f + 2) * 64] * v2;
    }
#pragma unroll
    for (int f = 0; f < NK_MAXF; f++) {
        if (f < nf) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) s[f] += __shfl_xor(s[f], o);
            if (l == 0) sh[w][f] = s[f];
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < nf) {
        double t = 0.0;
        for (int q = 0; q < nw; q++) t += sh[q][threadIdx.x];
        rc[I * nf + threadIdx.x] = t;
    }
}

// wavefronts of k_coarse_restrict: one per 64-point chunk of an aggregate, at most 16
static inline dim3 restrict_block(int agg) { return dim3((unsigned)(64 * std::min(agg / 64, 16))); }

// zc = Ainv rc: four rows per workgroup (a wavefront each, four independent partial sums per
// lane), rc staged in LDS (dynamic: mc doubles)
__global__ __launch_bounds__(256) void k_coarse_matvec(int64_t mc, const float *Ainv, const double *rc,
                                                      double *zc)
{
    extern __shared__ double rcs[];
    for (int64_t k = threadIdx.x; k < mc; k += 256) rcs[k] = rc[k];
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= mc) return;  // after the only barrier
    const int l = threadIdx.x & 63;
    const float *row = Ainv + r * mc;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int64_t k = l;
    for (; k + 192 < mc; k += 256) {
        s0 += (double)row[k] * rcs[k];
        s1 += (double)row[k + 64] * rcs[k + 64];
        s2 += (double)row[k + 128] * rcs[k + 128];
        s3 += (double)row[k + 192] * rcs[k + 192];
    }
    for (; k < mc; k += 64) s0 += (double)row[k] * rcs[k];
    double s = (s0 + s1) + (s2 + s3);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (l == 0) zc[r] = s;
}

// z = a D^-1 v + b P zc  (a = 0: the prolongation alone, added to z when b is given and
// `add` is set)
// partial sum of x over a 256-thread block into parts[blockIdx.x] (fixed order)
__device__ __forceinline__ void nk_block_part(double x, double *parts)
{
    __shared__ double sh[4];
    x = cgs_wave_sum(x);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) parts[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ __launch_bounds__(256) void k_apply_two_level(int64_t n, int agg, int nf, const double *dinv,
                                                         const double *v, const double *zc,
                                                         const double *phis, double alpha, int add,
                                                         double *z, double *parts)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double zz = 0.0;
    if (i < n) {
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
        const double *b = phis + (i >> 6) * nf * 3 * 64 + (i & 63);
        for (int f = 0; f < nf; f++) {
            const double w = alpha * zc[I * nf + f];
            o[0] += w * b[(3 * f) * 64];
            o[1] += w * b[(3 * f + 1) * 64];
            o[2] += w * b[(3 * f + 2) * 64];
        }
        for (int d = 0; d < 3; d++) z[3 * i + d] = o[d];
        zz = (o[0] * o[0] + o[1] * o[1]) + o[2] * o[2];
    }
    // every thread of the block reaches the partial sum (it holds a barrier)
    if (parts) nk_block_part(zz, parts);
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
        for (void *p : {(void *)k->phi, (void *)k->phis, (void *)k->Ycs, (void *)k->Ac, (void *)k->Ainv,
                        (void *)k->rc, (void *)k->zc})
            if (p) (void)hipFree(p);
        k->phi = k->phis = k->Ycs = k->Ac = k->rc = k->zc = nullptr;
        k->Ainv = nullptr;
        const int64_t mc = k->nagg * nf;
        PDX_TRY(pdx_alloc(c, &k->phi, (size_t)n * nf * 3));
        PDX_TRY(pdx_alloc(c, &k->phis, (size_t)((n + 63) / 64) * 64 * nf * 3));
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
    hipLaunchKernelGGL(k_coarse_soa, grid1((n + 63) / 64 * 64), dim3(256), 0, s, n, nf, k->phi, k->phis);
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
pdx_status pdx_nk_two_level(pdx_ctx *c, const double *v, double *z, double *parts)
{
    pdx_nk *k = c->nk;
    if (!k->coarse_ready) return pdx_fail(c, PDX_ERR_STATE, "coarse space not built");
    hipLaunchKernelGGL(k_coarse_restrict, dim3((unsigned)k->nagg), restrict_block(k->agg), 0, c->stream, c->n, k->agg,
                       k->nf, k->phis, v, k->rc);
    hipLaunchKernelGGL(k_coarse_matvec, dim3((unsigned)((k->mc + 3) / 4)), dim3(256), sizeof(double) * k->mc,
                       c->stream, k->mc, k->Ainv, k->rc, k->zc);
    hipLaunchKernelGGL(k_apply_two_level, grid1(c->n), dim3(256), 0, c->stream, c->n, k->agg, k->nf, k->dinv, v,
                       k->zc, k->phis, 1.0, 0, z, parts);
    PDX_HIP(c, hipGetLastError());
    return PDX_OK;
}

// z += alpha P A_c^-1 P^T v
pdx_status pdx_nk_coarse_add(pdx_ctx *c, const double *v, double *z, double alpha)
{
    pdx_nk *k = c->nk;
    if (!k->coarse_ready) return pdx_fail(c, PDX_ERR_STATE, "coarse space not built");
    hipLaunchKernelGGL(k_coarse_restrict, dim3((unsigned)k->nagg), restrict_block(k->agg), 0, c->stream, c->n, k->agg,
                       k->nf, k->phis, v, k->rc);
    hipLaunchKernelGGL(k_coarse_matvec, dim3((unsigned)((k->mc + 3) / 4)), dim3(256), sizeof(double) * k->mc,
                       c->stream, k->mc, k->Ainv, k->rc, k->zc);
    hipLaunchKernelGGL(k_apply_two_level, grid1(c->n), dim3(256), 0, c->stream, c->n, k->agg, k->nf, k->dinv, v,
                       k->zc, k->phis, alpha, 1, z, (double *)nullptr);
    PDX_HIP(c, hipGetLastError());
    return PDX_OK;
}

// z = M^-1 v for the chosen preconditioner

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

// ----------------------------------------------------------------------------------------
// Flexible GMRES(restart) for J du = b, du in k->du (starts at zero).  Stops at
// ||res|| <= max(reltol ||b||, abstol) or after maxiter products.  Classical Gram-Schmidt
// with one reorthogonalisation pass.  The Arnoldi coefficients, the Hessenberg matrix, its
// Givens rotations and the residual estimate stay on the device (k_arnoldi); the host
// keeps one iteration queued ahead and reads the convergence flag of the previous one, so
// no iteration waits for a round trip.  An iteration queued after convergence leaves the
// Hessenberg matrix alone and its direction unused.
// ----------------------------------------------------------------------------------------
// device layout of the small GMRES state (k->gm): H (R+1) x R column-major, cs, sn, g, y,
// then the integer flags
struct nk_gm_view {
    double *H, *cs, *sn, *g, *y, *h1, *h2, *hn;
    int *flags;  // [0] done, [1] columns of the cycle
};
static nk_gm_view nk_gm(pdx_nk *k, int R)
{
    nk_gm_view v;
    v.H = k->gm;
    v.cs = v.H + (size_t)(R + 1) * R;
    v.sn = v.cs + R;
    v.g = v.sn + R;
    v.y = v.g + (R + 1);
    v.h1 = v.y + R;
    v.h2 = v.h1 + (R + 1);
    v.hn = v.h2 + (R + 1);
    v.flags = reinterpret_cast<int *>(v.hn + 2);
    return v;
}

// The next basis vector and its preconditioned image; a block of 192 threads takes 64
// points, one dof per thread:
//   V = (w - V h) / hn  (hn_sq: / sqrt(hn), the cycle's first vector r / beta), in place
//   when vout = w; with dinv (block Jacobi) also z = D^-1 V (the point's three values
//   through LDS) and the block's partial sum of |z|^2 into parts (the J v of z takes ||z||
//   from them)
__global__ __launch_bounds__(192) void k_next_basis(int64_t n, int nv, const double *h, const double *hn,
                                                    int hn_sq, const double *V, int64_t ldv,
                                                    const double *w, double *vout, const double *dinv,
                                                    double *z, double *parts)
{
    __shared__ double xs[192];
    __shared__ double sh[3];
    const int64_t m = 3 * n;
    const int64_t t = (int64_t)blockIdx.x * 192 + threadIdx.x;
    double x = 0.0;
    if (t < m) {
        const double raw = *hn;
        const double dv = hn_sq ? sqrt(raw) : raw;
        double a = w[t];
#pragma unroll 8
        for (int k = 0; k < nv; k++) a -= h[k] * V[(int64_t)k * ldv + t];
        x = raw > 0.0 ? a / dv : a;
        vout[t] = x;
    }
    if (!dinv) return;  // uniform over the grid
    xs[threadIdx.x] = x;
    __syncthreads();
    double zz = 0.0;
    if (t < m) {
        const int pl = threadIdx.x / 3, d = threadIdx.x % 3;
        const double *R = dinv + 9 * (t / 3);
        const double x0 = xs[3 * pl], x1 = xs[3 * pl + 1], x2 = xs[3 * pl + 2];
        const double zd = R[d] * x0 + R[3 + d] * x1 + R[6 + d] * x2;
        z[t] = zd;
        zz = zd * zd;
    }
    zz = cgs_wave_sum(zz);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = zz;
    __syncthreads();
    if (threadIdx.x == 0) parts[blockIdx.x] = (sh[0] + sh[1]) + sh[2];
}

// start of a cycle: g = (beta, 0, ...), flags cleared; beta^2 in *b2
__global__ void k_gm_start(int R, const double *b2, double *g, int *flags)
{
    const int t = threadIdx.x;
    for (int i = t; i <= R; i += blockDim.x) g[i] = (i == 0) ? sqrt(*b2) : 0.0;
    if (t == 0) {
        flags[0] = 0;
        flags[1] = 0;
    }
}

// Column j of the Hessenberg matrix.  h1: the first Gram-Schmidt pass (reduced); the
// second pass arrives as j + 2 columns of np partial sums (V_i . w for i <= j, then n1 =
// |w|^2 after the first pass), reduced here into h2 (one wave per sum); since V is
// orthonormal, |w - V h2|^2 = n1 - |h2|^2.  The rotations run on the column staged in
// LDS.  Stores the new norm in hn[0], the columns so far in flags[1]; sets flags[0] (done)
// when the residual estimate |g[j + 1]| reaches thr, the norm vanishes or the iteration is
// the last one allowed, and writes 2 tok + flags[0] to flag_host[j % PDX_NK_SLOTS] (pinned host
// memory, may be null; tok tells the host which column wrote the slot).  One workgroup of
// 1024 threads.
__global__ __launch_bounds__(1024) void k_arnoldi(int j, int R, const double *h1, const double *parts,
                                                  int np, double *h2, double *hn, double thr,
                                                  int last_allowed, double *H, double *cs, double *sn,
                                                  double *g, int *flags, int *flag_host, int tok)
{
    __shared__ double h2s[PDX_MAX_RESTART + 2];
    __shared__ double col[PDX_MAX_RESTART + 2], c_[PDX_MAX_RESTART], s_[PDX_MAX_RESTART];
    __shared__ double hn_s;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int nres = j + 2;
    for (int q = wv; q < nres; q += 16) {
        double a = 0.0;
#pragma unroll 8
        for (int b = lane; b < np; b += 64) a += parts[(int64_t)q * np + b];
        a = cgs_wave_sum(a);
        if (lane == 0) {
            h2s[q] = a;
            h2[q] = a;
        }
    }
    __syncthreads();
    if (wv == 0) {
        double q2 = 0.0;
        for (int i = lane; i <= j; i += 64) q2 += h2s[i] * h2s[i];
        q2 = cgs_wave_sum(q2);
        if (lane == 0) {
            hn_s = sqrt(fmax(h2s[j + 1] - q2, 0.0));
            hn[0] = hn_s;
        }
    }
    const bool done_before = flags[0] != 0;
    for (int i = t; i <= j; i += 1024) {
        col[i] = h1[i] + h2s[i];
        if (i < j) {
            c_[i] = cs[i];
            s_[i] = sn[i];
        }
    }
    __syncthreads();
    if (t == 0) {
        const double hnv = hn_s;
        if (done_before) {
            if (flag_host) __hip_atomic_store(flag_host + j % PDX_NK_SLOTS, 2 * tok + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        } else {
            col[j + 1] = hnv;
            // the previous rotations, eight at a time: their inputs are read from LDS ahead
            // of the dependent chain (the running value a is carried in a register)
            double a = col[0];
            for (int i0 = 0; i0 < j; i0 += 8) {
                double cc[8], ss[8], bb[8];
#pragma unroll
                for (int u = 0; u < 8; u++)
                    if (i0 + u < j) {
                        cc[u] = c_[i0 + u];
                        ss[u] = s_[i0 + u];
                        bb[u] = col[i0 + u + 1];
                    }
#pragma unroll
                for (int u = 0; u < 8; u++)
                    if (i0 + u < j) {
                        col[i0 + u] = cc[u] * a + ss[u] * bb[u];
                        a = -ss[u] * a + cc[u] * bb[u];
                    }
            }
            col[j] = a;
            const double den = hypot(col[j], col[j + 1]);
            const double cj = den != 0 ? col[j] / den : 1.0;
            const double sj = den != 0 ? col[j + 1] / den : 0.0;
            cs[j] = cj;
            sn[j] = sj;
            col[j] = den;
            col[j + 1] = 0.0;
            const double gj = g[j];
            g[j + 1] = -sj * gj;
            g[j] = cj * gj;
            flags[1] = j + 1;
            const int done = (fabs(-sj * gj) <= thr || last_allowed || hnv == 0) ? 1 : 0;
            flags[0] = done;
            if (flag_host) __hip_atomic_store(flag_host + j % PDX_NK_SLOTS, 2 * tok + done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    __syncthreads();
    if (done_before) return;
    double *Hc = H + (size_t)(R + 1) * j;
    for (int i = t; i <= j + 1; i += 1024) Hc[i] = col[i];
}

// y = H^-1 g for the cycle's columns (upper triangular, column-oriented back substitution on
// one wave: lane l keeps rows l and l + 64), zero beyond; est = |g[cols]|
__global__ __launch_bounds__(64) void k_gm_solve(int R, const double *H, const double *g,
                                                 const int *flags, double *y, double *est)
{
    const int t = threadIdx.x;
    const int j = flags[1];
    const size_t ld = (size_t)(R + 1);
    double r0 = t < j ? g[t] : 0.0, r1 = t + 64 < j ? g[t + 64] : 0.0;
    double y0 = 0.0, y1 = 0.0;
    for (int i = j - 1; i >= 0; i--) {
        const double *Hi = H + ld * i;
        // lane (i & 63) owns row i
        double yi = 0.0;
        if (t == (i & 63)) yi = (i < 64 ? r0 : r1) / Hi[i];
        yi = __shfl(yi, i & 63);
        if (t == (i & 63)) {
            if (i < 64) y0 = yi;
            else y1 = yi;
        }
        if (t < i) r0 -= Hi[t] * yi;
        if (t + 64 < i) r1 -= Hi[t + 64] * yi;
    }
    if (t < R) y[t] = y0;
    if (t + 64 < R) y[t + 64] = y1;
    if (t == 0) *est = fabs(g[j]);
}

// Waits until the Arnoldi kernel of the column with token tok has written its flag into slot
// (2 tok + done, a system-scope store into pinned memory) by spinning on that word, with no
// event in the stream; after 2 s without it, synchronises the stream and reads it once more.
static pdx_status nk_wait_flag(pdx_ctx *c, int slot, int tok, int *done)
{
    pdx_nk *k = c->nk;
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t spin = 1;; spin++) {
        const int v = __atomic_load_n(k->flag_host + slot, __ATOMIC_ACQUIRE);
        if ((v >> 1) == tok) {
            *done = v & 1;
            return PDX_OK;
        }
        if ((spin & 1023) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2)) {
            PDX_HIP(c, hipStreamSynchronize(c->stream));
            const int w = __atomic_load_n(k->flag_host + slot, __ATOMIC_ACQUIRE);
            if ((w >> 1) != tok) return pdx_fail(c, PDX_ERR_HIP, "a Krylov column's flag was not written");
            *done = w & 1;
            return PDX_OK;
        }
    }
}

// |x|^2 into out_dev
static pdx_status nk_norm2_dev(pdx_ctx *c, const double *x, double *out_dev)
{
    pdx_nk *k = c->nk;
    const int nb = cgs_blocks(k->m);
    hipLaunchKernelGGL(k_cgs_dots, dim3(nb), dim3(CGS_T), 0, c->stream, k->m, 1, x, k->m, x, k->red);
    hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, c->stream, nb, (const double *)k->red, out_dev);
    PDX_HIP(c, hipGetLastError());
    return PDX_OK;
}

// The next basis vector V = (w - V h) / hn (or r / sqrt(hn), hn_sq = 1) into vout and, when
// zout is given, its preconditioned image z = M^-1 V with the partial sums of |z|^2 (*parts,
// *np; null where the preconditioner leaves none).  Block Jacobi is one kernel; the
// two-level (translations or learned fields) adds its coarse correction.
static pdx_status nk_next(pdx_ctx *c, int precond, int nv, const double *h, const double *hn, int hn_sq,
                          const double *w, double *vout, double *zout, const double **parts, int *np)
{
    pdx_nk *k = c->nk;
    const int64_t n = c->n;
    const bool jac = zout && precond == PDX_PRE_JACOBI;
    const unsigned nbb = (unsigned)((n + 63) / 64);
    hipLaunchKernelGGL(k_next_basis, dim3(nbb), dim3(192), 0, c->stream, n, nv, h, hn, hn_sq,
                       (const double *)k->Vb, k->m, w, vout, jac ? (const double *)k->dinv : nullptr,
                       jac ? zout : nullptr, jac ? k->zparts : nullptr);
    PDX_HIP(c, hipGetLastError());
    *parts = nullptr;
    *np = 0;
    if (!zout) return PDX_OK;
    if (precond == PDX_PRE_TWO_LEVEL || precond == PDX_PRE_LEARNED)
        PDX_TRY(pdx_nk_two_level(c, vout, zout, k->zparts));  // learned: the coarse space holds its fields
    else if (!jac)
        PDX_HIP(c, hipMemcpyAsync(zout, vout, sizeof(double) * k->m, hipMemcpyDeviceToDevice, c->stream));
    if (jac) {
        *parts = k->zparts;
        *np = (int)nbb;
    } else if (precond == PDX_PRE_TWO_LEVEL || precond == PDX_PRE_LEARNED) {
        *parts = k->zparts;
        *np = (int)grid1(n).x;
    }
    return PDX_OK;
}

static pdx_status fgmres(pdx_ctx *c, const double *b, int restart, int maxiter, double reltol,
                         double abstol, int precond, double scale, int *iters_out,
                         double *res_out)
{
    pdx_nk *k = c->nk;
    const int64_t m = k->m;
    const int R = restart;
    nk_gm_view G = nk_gm(k, R);
    hipStream_t s = c->stream;
    const int CB = cgs_blocks(m);
    PDX_HIP(c, hipMemsetAsync(k->du, 0, sizeof(double) * m, s));
    PDX_TRY(nk_refresh_unorm(c));
    // |b|^2 -> the cycle's beta^2 (G.hn holds it until the first column)
    PDX_TRY(nk_norm2_dev(c, b, G.hn));
    double bn2 = 0.0;
    PDX_HIP(c, hipMemcpyAsync(&bn2, G.hn, sizeof(double), hipMemcpyDeviceToHost, s));
    PDX_HIP(c, hipStreamSynchronize(s));
    const double bnorm = std::sqrt(bn2);
    const double thr = std::max(reltol * bnorm, abstol);
    int it = 0;
    PDX_HIP(c, hipMemcpyAsync(k->r, b, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
    double beta = bnorm;
    while (it < maxiter && beta > thr) {
        hipLaunchKernelGGL(k_gm_start, dim3(1), dim3(64), 0, s, R, G.hn, G.g, G.flags);
        PDX_HIP(c, hipGetLastError());
        // V_0 = r / beta and Z_0 = M^-1 V_0
        const double *zp = nullptr;
        int znp = 0;
        PDX_TRY(nk_next(c, precond, 0, nullptr, G.hn, 1, k->r, k->Vb, k->Zb, &zp, &znp));
        int enq = 0;  // columns queued in this cycle
        for (int j = 0; j < R && it + enq < maxiter; j++) {
            double *Zj = k->Zb + (int64_t)j * m;
            double *w = k->Vb + (int64_t)(j + 1) * m;
            PDX_TRY(pdx_nk_jv_dev(c, Zj, w, scale, zp, znp));
            if (pdx_gnn_collecting(c)) PDX_TRY(pdx_gnn_record_krylov(c, Zj, w));
            // classical Gram-Schmidt, twice: h1 = V^T w; w -= V h1, then V^T w and |w|^2 in
            // partial sums that the Arnoldi step reduces into h2; then the next basis vector
            // (w - V h2) / |w - V h2| and its preconditioned image
            const int nv = j + 1;
            hipLaunchKernelGGL(k_cgs_dots, dim3(CB), dim3(CGS_T), 0, s, m, nv, (const double *)k->Vb, m,
                               (const double *)w, k->red);
            hipLaunchKernelGGL(k_reduce_cols, dim3((unsigned)nv), dim3(256), 0, s, CB, (const double *)k->red, G.h1);
            if (nv <= CGS_REG && m <= (int64_t)CB * CGS_T)
                hipLaunchKernelGGL(k_cgs_axdn, dim3(CB), dim3(CGS_T), 0, s, m, nv, (const double *)G.h1,
                                   (const double *)k->Vb, m, w, k->red);
            else
                hipLaunchKernelGGL(k_cgs_axdn_loop, dim3(CB), dim3(CGS_T), 0, s, m, nv, (const double *)G.h1,
                                   (const double *)k->Vb, m, w, k->red);
            const int last_allowed = (it + enq + 1 >= maxiter) ? 1 : 0;
            const int tok = k->col_tok = k->col_tok % 0x3fffffff + 1;  // 1 .. 2^30 - 1: 2 tok + 1 fits an int
            k->tok[j % PDX_NK_SLOTS] = tok;
            hipLaunchKernelGGL(k_arnoldi, dim3(1), dim3(1024), 0, s, j, R, (const double *)G.h1,
                               (const double *)k->red, CB, G.h2, G.hn, thr, last_allowed, G.H, G.cs, G.sn,
                               G.g, G.flags, k->flag_dev, tok);
            PDX_HIP(c, hipGetLastError());
            // Z has R columns: the last column's V needs no image
            PDX_TRY(nk_next(c, precond, nv, G.h2, G.hn, 0, w, w, j + 1 < R ? k->Zb + (int64_t)(j + 1) * m : nullptr,
                            &zp, &znp));
            enq++;
            // with this column queued, wait for the flag of the column PDX_NK_LAG before it
            if (j >= PDX_NK_LAG) {
                const int jw = j - PDX_NK_LAG;
                int done = 0;
                PDX_TRY(nk_wait_flag(c, jw % PDX_NK_SLOTS, k->tok[jw % PDX_NK_SLOTS], &done));
                if (done) break;
            }
        }
        hipLaunchKernelGGL(k_gm_solve, dim3(1), dim3(64), 0, s, R, G.H, G.g, G.flags, G.y, G.hn + 1);
        hipLaunchKernelGGL(k_lincomb_add, grid1(m), dim3(256), 0, s, m, enq, k->Zb, m, G.y, k->du);
        PDX_HIP(c, hipGetLastError());
        int cols = 0;
        double est = 0.0;
        PDX_HIP(c, hipMemcpyAsync(&cols, G.flags + 1, sizeof(int), hipMemcpyDeviceToHost, s));
        PDX_HIP(c, hipMemcpyAsync(&est, G.hn + 1, sizeof(double), hipMemcpyDeviceToHost, s));
        PDX_HIP(c, hipStreamSynchronize(s));
        it += cols;
        if (est <= thr || it >= maxiter || cols == 0) {
            beta = est;
            break;
        }
        // true residual for the restart: r = b - J du
        PDX_TRY(pdx_nk_jv_dev(c, k->du, k->r, scale));
        it++;
        hipLaunchKernelGGL(k_sub_from, grid1(m), dim3(256), 0, s, m, b, k->r);
        PDX_HIP(c, hipGetLastError());
        PDX_TRY(nk_norm2_dev(c, k->r, G.hn));
        double r2 = 0.0;
        PDX_HIP(c, hipMemcpyAsync(&r2, G.hn, sizeof(double), hipMemcpyDeviceToHost, s));
        PDX_HIP(c, hipStreamSynchronize(s));
        beta = std::sqrt(r2);
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
