// Adaptive dynamic relaxation on the device (src/time_solvers/dynamic_relaxation.jl).
//
// One step n (relaxation_timestep!, :219-240):
//   velocity BCs on v_half at t = n dt, u += v_half dt, x += v_half dt   (k_dr_begin)
//   force density at x                                                  (the step kernel,
//                                                                        force only)
//   n = 1: v_half = dt/2 (b_int + b_ext) / Lambda  (relaxation_first_step!, :326-333)
//   n > 1: cn = 2 sqrt(cn1 / cn2), clamped (calc_damping, :262-323), with
//          cn1 = sum over dofs with v_half != 0 of -u^2 (b_int - b_int_old) /
//                (dt Lambda v_half_old),  cn2 = sum u^2, both compensated (two_sum);
//          v_half = ((2 - cn dt) v_half + 2 dt a) / (2 + cn dt), a = (b_int + b_ext) / Lambda
//          (relaxation_step!, :335-343)
//   v = (v_half_old + v_half) / 2, v_half_old = v_half, b_int_old = b_int
//   (relaxation_updates!, :345-350)
// Lambda per point (calc_density_matrix!, :166-176):
//   damping_factor * 1 / 4 * dt^2 * n_bonds * (5 pi delta^2 bc), the same for x, y, z.
#include <cmath>

#include "pdx_internal.h"

#define DR_BLOCKS 1024

__global__ void k_dr_density(int64_t n, const int32_t *cnt, double lam, double dt, double delta,
                             double bc, const pdx_pset *ps, const uint8_t *pid, double *density)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (ps) bc = ps[pid[i]].bc;  // the point's own set (get_params(paramsetup, i))
    const double k = ((5.0 * M_PI) * (delta * delta)) * bc;
    density[i] = ((((lam * 1.0) / 4.0) * (dt * dt)) * (double)cnt[i]) * k;
}

// velocity BCs on v_half (apply_boundary_conditions!), then u and x
__global__ void k_dr_begin(int64_t n, const int32_t *bcslot, const int32_t *slot_src,
                           const double *bcv, double dt, double *vh, double *u, double4 *rec)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * n) return;
    const int64_t i = t / 3;
    const int d = (int)(t - 3 * i);
    double v = vh[t];
    const int32_t s = bcslot[i];
    if (s >= 0 && bcv) {
        const int32_t src = slot_src[3 * s + d];
        if (src >= 0) {
            const double val = bcv[src];
            if (isfinite(val)) v = val;
        }
    }
    vh[t] = v;
    const double du = v * dt;
    u[t] = u[t] + du;
    double *x = reinterpret_cast<double *>(rec) + 4 * i + d;
    *x = *x + du;
}

__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e)
{
    s = a + b;
    const double v = s - a;
    e = (a - (s - v)) + (b - v);
}

// compensated partial sums of cn1 and cn2: part[4 blk + (cn1, cn1_c, cn2, cn2_c)]
__global__ __launch_bounds__(256) void k_dr_damping(int64_t m, const double *u, const double *vh,
                                                    const double *vh_old, const double *bint,
                                                    const double *bint_old, const double *density,
                                                    double dt, double *part)
{
    __shared__ double sh[4][256];
    double c1 = 0.0, c1c = 0.0, c2 = 0.0, c2c = 0.0, e;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < m; t += (int64_t)gridDim.x * 256) {
        const double ut = u[t];
        if (vh[t] != 0.0) {
            const double db = bint[t] - bint_old[t];
            const double temp = (dt * density[t / 3]) * vh_old[t];
            two_sum(c1, ((-(ut * ut)) * db) / temp, c1, e);
            c1c += e;
        }
        two_sum(c2, ut * ut, c2, e);
        c2c += e;
    }
    sh[0][threadIdx.x] = c1;
    sh[1][threadIdx.x] = c1c;
    sh[2][threadIdx.x] = c2;
    sh[3][threadIdx.x] = c2c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            double s;
            two_sum(sh[0][threadIdx.x], sh[0][threadIdx.x + o], s, e);
            sh[0][threadIdx.x] = s;
            sh[1][threadIdx.x] += sh[1][threadIdx.x + o] + e;
            two_sum(sh[2][threadIdx.x], sh[2][threadIdx.x + o], s, e);
            sh[2][threadIdx.x] = s;
            sh[3][threadIdx.x] += sh[3][threadIdx.x + o] + e;
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// combine the partial sums (two_sum) and finish_damping: cn into out[0]
__global__ __launch_bounds__(256) void k_dr_finish(int nblk, const double *part, double *out)
{
    __shared__ double sh[4][256];
    double c1 = 0.0, c1c = 0.0, c2 = 0.0, c2c = 0.0, s, e;
    for (int b = threadIdx.x; b < nblk; b += 256) {
        two_sum(c1, part[4 * b], s, e);
        c1 = s;
        c1c += part[4 * b + 1] + e;
        two_sum(c2, part[4 * b + 2], s, e);
        c2 = s;
        c2c += part[4 * b + 3] + e;
    }
    sh[0][threadIdx.x] = c1;
    sh[1][threadIdx.x] = c1c;
    sh[2][threadIdx.x] = c2;
    sh[3][threadIdx.x] = c2c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            two_sum(sh[0][threadIdx.x], sh[0][threadIdx.x + o], s, e);
            sh[0][threadIdx.x] = s;
            sh[1][threadIdx.x] += sh[1][threadIdx.x + o] + e;
            two_sum(sh[2][threadIdx.x], sh[2][threadIdx.x + o], s, e);
            sh[2][threadIdx.x] = s;
            sh[3][threadIdx.x] += sh[3][threadIdx.x + o] + e;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double cn1 = sh[0][0] + sh[1][0], cn2 = sh[2][0] + sh[3][0];
        double cn = 0.0;
        if (cn2 != 0.0 && cn1 / cn2 > 0.0) cn = 2.0 * sqrt(cn1 / cn2);
        if (cn > 2.0) cn = 1.9;
        out[0] = cn;
    }
}

// relaxation_first_step! / relaxation_step! and relaxation_updates!
__global__ void k_dr_update(int64_t m, int first, double dt, const double *cnp, const double *bint,
                            const double *bext, const double *density, double *vh, double *vh_old,
                            double *v, double *bint_old)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const double rho = density[t / 3];
    const double bt = bint[t] + (bext ? bext[t] : 0.0);
    double vn;
    if (first) {
        vn = ((0.5 * dt) * bt) / rho;
    } else {
        const double cn = *cnp;
        const double a = bt / rho;
        vn = ((2.0 - cn * dt) * vh[t] + (2.0 * dt) * a) / (2.0 + cn * dt);
    }
    vh[t] = vn;
    v[t] = 0.5 * (vh_old[t] + vn);
    vh_old[t] = vn;
    bint_old[t] = bint[t];
}

static pdx_status dr_ensure(pdx_ctx *c)
{
    if (c->dr_vh_old) return PDX_OK;
    const int64_t m = 3 * c->n;
    PDX_TRY(pdx_alloc(c, &c->dr_vh_old, m));
    PDX_TRY(pdx_alloc(c, &c->dr_bint_old, m));
    PDX_TRY(pdx_alloc(c, &c->dr_density, c->n));
    PDX_TRY(pdx_alloc(c, &c->dr_part, 4 * DR_BLOCKS + 8));
    PDX_HIP(c, hipMemsetAsync(c->dr_vh_old, 0, sizeof(double) * m, c->stream));
    PDX_HIP(c, hipMemsetAsync(c->dr_bint_old, 0, sizeof(double) * m, c->stream));
    PDX_HIP(c, hipMemsetAsync(c->dr_density, 0, sizeof(double) * c->n, c->stream));
    return PDX_OK;
}

static inline dim3 grid_of(int64_t m) { return dim3((unsigned)((m + 255) / 256)); }

extern "C" pdx_status pdx_dr_init(pdx_ctx *c, double dt, double damping_factor)
{
    if (!c) return PDX_ERR_ARG;
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    if (!(dt > 0)) return pdx_fail(c, PDX_ERR_ARG, "`stepsize` should be larger than zero!");
    if (!(damping_factor > 0)) return pdx_fail(c, PDX_ERR_ARG, "`damping_factor` should be larger than zero!");
    PDX_HIP(c, hipSetDevice(c->device));
    PDX_TRY(dr_ensure(c));
    hipLaunchKernelGGL(k_dr_density, grid_of(c->n), dim3(256), 0, c->stream, c->n, c->cnt, damping_factor,
                       dt, c->P.delta, c->P.bc, c->multi() ? c->d_psets : nullptr, c->d_pid, c->dr_density);
    PDX_HIP(c, hipGetLastError());
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    c->dr_ready = true;
    return PDX_OK;
}

extern "C" pdx_status pdx_dr_damping(pdx_ctx *c, double dt, double *cn_out)
{
    if (!c || !cn_out) return PDX_ERR_ARG;
    if (!c->dr_ready) return pdx_fail(c, PDX_ERR_STATE, "pdx_dr_init first");
    PDX_HIP(c, hipSetDevice(c->device));
    const int64_t m = 3 * c->n;
    const int nb = (int)std::min<int64_t>(DR_BLOCKS, (m + 255) / 256);
    hipLaunchKernelGGL(k_dr_damping, dim3(nb), dim3(256), 0, c->stream, m, c->u, c->vh, c->dr_vh_old,
                       c->bint, c->dr_bint_old, c->dr_density, dt, c->dr_part);
    hipLaunchKernelGGL(k_dr_finish, dim3(1), dim3(256), 0, c->stream, nb, c->dr_part,
                       c->dr_part + 4 * DR_BLOCKS);
    PDX_HIP(c, hipGetLastError());
    PDX_HIP(c, hipMemcpyAsync(cn_out, c->dr_part + 4 * DR_BLOCKS, sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    return PDX_OK;
}

extern "C" pdx_status pdx_dr_run(pdx_ctx *c, int64_t first, int64_t n_steps, double dt,
                                 const double *bc_values, int64_t *nan_step_out)
{
    if (!c) return PDX_ERR_ARG;
    if (nan_step_out) *nan_step_out = -1;
    if (!c->dr_ready) return pdx_fail(c, PDX_ERR_STATE, "pdx_dr_init first");
    if (n_steps <= 0) return PDX_OK;
    if (first < 1) return pdx_fail(c, PDX_ERR_ARG, "steps count from 1");
    if (!(dt > 0)) return pdx_fail(c, PDX_ERR_ARG, "`stepsize` should be larger than zero!");
    if (c->n_src > 0 && !bc_values) return pdx_fail(c, PDX_ERR_ARG, "missing BC values");
    PDX_HIP(c, hipSetDevice(c->device));
    const int64_t need = std::max<int64_t>(c->n_src * n_steps, 1);
    if (need > c->bcvals_cap) {
        if (c->bcvals) (void)hipFree(c->bcvals);
        c->bcvals = nullptr;
        PDX_TRY(pdx_alloc(c, &c->bcvals, need));
        c->bcvals_cap = need;
    }
    if (c->n_src > 0)
        PDX_HIP(c, hipMemcpyAsync(c->bcvals, bc_values, sizeof(double) * c->n_src * n_steps,
                                  hipMemcpyHostToDevice, c->stream));
    PDX_TRY(pdx_reset_flags(c));
    const int64_t m = 3 * c->n;
    const int nb = (int)std::min<int64_t>(DR_BLOCKS, (m + 255) / 256);
    double *cnp = c->dr_part + 4 * DR_BLOCKS;
    for (int64_t k = 0; k < n_steps; k++) {
        const int64_t step = first + k;
        const double *bcv = c->n_src > 0 ? c->bcvals + c->n_src * k : nullptr;
        hipLaunchKernelGGL(k_dr_begin, grid_of(m), dim3(256), 0, c->stream, c->n, c->bcslot, c->slot_src,
                           bcv, dt, c->vh, c->u, c->rec[c->cur]);
        PDX_HIP(c, hipGetLastError());
        PDX_TRY(pdx_apply_fd_bcs(c, bcv));
        pdx_step_args a;
        PDX_TRY(pdx_fill_step_args(c, &a, step, PDX_F_FORCE_ONLY, dt, nullptr));
        PDX_TRY(pdx_launch_step(c, a));
        const int firstflag = step == 1 ? 1 : 0;
        if (!firstflag) {
            hipLaunchKernelGGL(k_dr_damping, dim3(nb), dim3(256), 0, c->stream, m, c->u, c->vh,
                               c->dr_vh_old, c->bint, c->dr_bint_old, c->dr_density, dt, c->dr_part);
            hipLaunchKernelGGL(k_dr_finish, dim3(1), dim3(256), 0, c->stream, nb, c->dr_part, cnp);
        }
        hipLaunchKernelGGL(k_dr_update, grid_of(m), dim3(256), 0, c->stream, m, firstflag, dt, cnp,
                           c->bint, c->bext_nonzero ? c->bext : nullptr, c->dr_density, c->vh,
                           c->dr_vh_old, c->v, c->dr_bint_old);
        PDX_HIP(c, hipGetLastError());
    }
    int64_t nan_step = -1;
    PDX_TRY(pdx_check_flags(c, &nan_step));
    if (nan_step >= 0) {
        if (nan_step_out) *nan_step_out = nan_step;
        return pdx_fail(c, PDX_ERR_NAN, "NaN in the force density at step %lld", (long long)nan_step);
    }
    return PDX_OK;
}
