// EnergySurfaceCorrection on the device (src/discretization/bond_system_corrections.jl:46-204,
// Le and Bobaru 2018).  A one-time pass after the bond search:
//
//   mfactor (calc_mfactor!, :98-122): for each direction d the body is stretched uniformly,
//     x = X with x_d = 1.001 X_d, and every point compares the strain energy density of
//     linear elasticity under that stretch, (lambda / 2 + mu) (1.001 - 1)^2
//     (stendens_uniext_small_strain, :147-149), with its peridynamic one
//     (strain_energy_density_point!, bond_based.jl:183-201, ordinary_state_based.jl:207-235)
//     at scfactor = 1: mfactor[d, i] = psi_theory / psi_pd.
//   scfactor (calc_scfactor!, :161-204): per bond, the ellipsoid radius of the two ends'
//     mean mfactors along the bond's direction angles theta, phi.
//
// One thread per point walks its SELL-64 row; the sums run in the device row order
// (Morton order of the neighbours), so they agree with the reference's bond order to
// rounding.  The step kernels then multiply every bond's weight by scfactor.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pdx_internal.h"

#define ESC_STRETCH 1.001

__device__ __forceinline__ int64_t esc_bond0(const int64_t *slice_ptr, int64_t p)
{
    return slice_ptr[p >> 6] + (p & 63);
}

// lambda, mu, bc, G, K of point i (several sets or one)
struct esc_par {
    double lam, mu, bc, G, K;
};

__device__ __forceinline__ esc_par esc_params(int64_t i, const esc_par &one, const pdx_pset *ps,
                                              const uint8_t *pid)
{
    if (!ps) return one;
    const pdx_pset &q = ps[pid[i]];
    return {q.lam, q.mu, q.bc, q.G, q.K};
}

__device__ __forceinline__ double3 esc_stretched(const double4 &X, int d)
{
    return make_double3(d == 0 ? ESC_STRETCH * X.x : X.x, d == 1 ? ESC_STRETCH * X.y : X.y,
                        d == 2 ? ESC_STRETCH * X.z : X.z);
}

// the OSB bond weight kernel(system, bond_id) (ordinary_state_based.jl:225): linear_kernel
// delta / L as Julia divides, the other influence functions as the step kernels evaluate them
__device__ __forceinline__ double esc_omega(int kid, double delta, double L)
{
    return kid == PDX_KERNEL_LINEAR ? delta / L : pdx_omega_other(kid, delta, L);
}

__global__ void k_esc_mfactor(int64_t n, int d, int32_t kind, int kid, const int64_t *slice_ptr,
                              const int32_t *cnt, const uint32_t *bw, const double *L,
                              const double4 *Xr, double delta, esc_par one, const pdx_pset *ps,
                              const uint8_t *pid, double *mfactor)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = cnt[i];
    const int64_t b0 = esc_bond0(slice_ptr, i);
    const esc_par Pi = esc_params(i, one, ps, pid);
    const double4 Xi = Xr[i];
    const double3 xi = esc_stretched(Xi, d);
    // get_averaged_lame_parameters (:124-145): the point's own (lambda, mu) with one set, the
    // mean over active bonds of the two ends' means with several
    double la = Pi.lam, mu = Pi.mu;
    if (ps) {
        double sl = 0.0, sm = 0.0;
        int na = 0;
        for (int t = 0; t < c; t++) {
            const uint32_t wd = bw[b0 + (int64_t)t * 64];
            if (!(wd & PDX_BW_ACT)) continue;
            const esc_par Pj = esc_params(wd & PDX_BW_NBR, one, ps, pid);
            sl += (Pi.lam + Pj.lam) / 2.0;
            sm += (Pi.mu + Pj.mu) / 2.0;
            na++;
        }
        la = sl / (double)na;
        mu = sm / (double)na;
    }
    const double st = ESC_STRETCH - 1.0;
    const double psi_t = (0.5 * la + mu) * (st * st);
    double psi = 0.0;
    if (kind == PDX_BB) {
        for (int t = 0; t < c; t++) {
            const int64_t b = b0 + (int64_t)t * 64;
            const uint32_t wd = bw[b];
            const int64_t j = wd & PDX_BW_NBR;
            const double4 Xj = Xr[j];
            const double3 xj = esc_stretched(Xj, d);
            const double dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
            const double l = sqrt((dx * dx + dy * dy) + dz * dz);
            const double Lb = L[b];
            const double eps = (l - Lb) / Lb;
            const double w = (wd & PDX_BW_ACT) ? 1.0 : 0.0;
            const double bcb = (Pi.bc + esc_params(j, one, ps, pid).bc) / 2.0;
            psi += (((((0.25 * w) * bcb) * eps) * eps) * Lb) * Xj.w;
        }
    } else {  // PDX_OSB
        double wvol = 0.0;
        for (int t = 0; t < c; t++) {
            const int64_t b = b0 + (int64_t)t * 64;
            const uint32_t wd = bw[b];
            const double4 Xj = Xr[wd & PDX_BW_NBR];
            const double Dx = Xj.x - Xi.x, Dy = Xj.y - Xi.y, Dz = Xj.z - Xi.z;
            const double w = (wd & PDX_BW_ACT) ? esc_omega(kid, delta, L[b]) : 0.0;
            wvol += ((w * 1.0) * ((Dx * Dx + Dy * Dy) + Dz * Dz)) * Xj.w;
        }
        if (wvol != 0.0) {  // a point without weighted volume keeps psi = 0
            const double c1 = 3.0 / wvol;
            double dil = 0.0;
            for (int t = 0; t < c; t++) {
                const int64_t b = b0 + (int64_t)t * 64;
                const uint32_t wd = bw[b];
                const double4 Xj = Xr[wd & PDX_BW_NBR];
                const double3 xj = esc_stretched(Xj, d);
                const double dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
                const double l = sqrt((dx * dx + dy * dy) + dz * dz);
                const double Lb = L[b];
                const double w = (wd & PDX_BW_ACT) ? esc_omega(kid, delta, Lb) : 0.0;
                dil += ((((w * 1.0) * c1) * Lb) * (l - Lb)) * Xj.w;
            }
            const double psi_vol = (0.5 * Pi.K) * (dil * dil);
            double psi_dev = 0.0;
            for (int t = 0; t < c; t++) {
                const int64_t b = b0 + (int64_t)t * 64;
                const uint32_t wd = bw[b];
                const int64_t j = wd & PDX_BW_NBR;
                const double4 Xj = Xr[j];
                const double3 xj = esc_stretched(Xj, d);
                const double dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
                const double l = sqrt((dx * dx + dy * dy) + dz * dz);
                const double Lb = L[b];
                const double e = l - Lb;
                const double edev = e - ((1.0 / 3.0) * dil) * Lb;
                const double w = (wd & PDX_BW_ACT) ? esc_omega(kid, delta, Lb) : 0.0;
                const double Gb = (Pi.G + esc_params(j, one, ps, pid).G) / 2.0;
                const double cdev = (15.0 * Gb) / (2.0 * wvol);
                psi_dev += ((((cdev * w) * 1.0) * edev) * edev) * Xj.w;
            }
            psi = psi_vol + psi_dev;
        }
    }
    mfactor[(int64_t)d * n + i] = psi_t / psi;
}

__global__ void k_esc_scfactor(int64_t n, const int64_t *slice_ptr, const int32_t *cnt,
                               const uint32_t *bw, const double *L, const double4 *Xr,
                               const double *mfactor, double *scf, float *scf32)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = cnt[i];
    const int64_t b0 = esc_bond0(slice_ptr, i);
    const double4 Xi = Xr[i];
    const double right = (90.0 * M_PI) / 180.0;
    for (int t = 0; t < c; t++) {
        const int64_t b = b0 + (int64_t)t * 64;
        const int64_t j = bw[b] & PDX_BW_NBR;
        const double4 Xj = Xr[j];
        const double ax = fabs(Xj.x - Xi.x), ay = fabs(Xj.y - Xi.y), az = fabs(Xj.z - Xi.z);
        const double scx = (mfactor[i] + mfactor[j]) / 2.0;
        const double scy = (mfactor[n + i] + mfactor[n + j]) / 2.0;
        const double scz = (mfactor[2 * n + i] + mfactor[2 * n + j]) / 2.0;
        double scr;
        if (az <= 1e-10 || !(ax <= 1e-10 && ay <= 1e-10)) {
            double th, ph;
            if (az <= 1e-10) {
                th = ay <= 1e-10 ? 0.0 : (ax <= 1e-10 ? right : atan(ay / ax));
                ph = right;
            } else {
                th = atan(ay / ax);
                ph = acos(az / L[b]);
            }
            const double cx = cos(th) * sin(ph), cy = sin(th) * sin(ph), cz = cos(ph);
            scr = sqrt(1.0 / (((cx * cx) / (scx * scx) + (cy * cy) / (scy * scy)) + (cz * cz) / (scz * scz)));
        } else {
            scr = scz;  // a bond along z
        }
        scf[b] = scr;
        scf32[b] = (float)scr;
    }
}

extern "C" pdx_status pdx_set_surface_correction(pdx_ctx *c, int32_t kind)
{
    if (!c) return PDX_ERR_ARG;
    if (c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "the surface correction must precede pdx_build_bonds");
    if (kind != 0 && kind != 1) return pdx_fail(c, PDX_ERR_ARG, "unknown surface correction %d", kind);
    c->corr = kind;
    return PDX_OK;
}

// called by pdx_build_bonds once the bonds, their states and the parameter sets are on the device
pdx_status pdx_build_correction(pdx_ctx *c)
{
    if (!c->corr) return PDX_OK;
    if (c->kind != PDX_BB && c->kind != PDX_OSB)
        return pdx_fail(c, PDX_ERR_UNSUPPORTED, "EnergySurfaceCorrection is built for BBMaterial and OSBMaterial");
    const int64_t n = c->n, nba = c->nb_alloc;
    PDX_TRY(pdx_alloc(c, &c->scf, (size_t)nba));
    PDX_TRY(pdx_alloc(c, &c->scf32, (size_t)nba));
    PDX_TRY(pdx_alloc(c, &c->mfactor, (size_t)(3 * n)));
    PDX_HIP(c, hipMemsetAsync(c->scf, 0, sizeof(double) * nba, c->stream));
    PDX_HIP(c, hipMemsetAsync(c->scf32, 0, sizeof(float) * nba, c->stream));
    const esc_par one = {c->P.lambda, c->P.mu, c->P.bc, c->P.G, c->P.K};
    const dim3 G((unsigned)((n + 255) / 256)), B(256);
    for (int d = 0; d < 3; d++)
        hipLaunchKernelGGL(k_esc_mfactor, G, B, 0, c->stream, n, d, c->kind, (int)c->P.kernel, c->slice_ptr,
                           c->cnt, c->bw,
                           c->L, c->Xr, c->P.delta, one, c->multi() ? c->d_psets : nullptr,
                           c->multi() ? c->d_pid : nullptr, c->mfactor);
    PDX_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_esc_scfactor, G, B, 0, c->stream, n, c->slice_ptr, c->cnt, c->bw, c->L, c->Xr,
                       c->mfactor, c->scf, c->scf32);
    PDX_HIP(c, hipGetLastError());
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    return PDX_OK;
}

extern "C" pdx_status pdx_get_surface_correction(pdx_ctx *c, double *scfactor, double *mfactor)
{
    if (!c) return PDX_ERR_ARG;
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    if (!c->corr) return pdx_fail(c, PDX_ERR_STATE, "no surface correction (pdx_set_surface_correction)");
    const int64_t n = c->n, nba = c->nb_alloc;
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    if (mfactor) {
        std::vector<double> m(3 * n);
        PDX_HIP(c, hipMemcpy(m.data(), c->mfactor, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; i++)
            for (int d = 0; d < 3; d++) mfactor[3 * i + d] = m[d * n + c->iperm[i]];
    }
    if (scfactor) {
        std::vector<int64_t> slots;
        PDX_TRY(pdx_caller_bond_slots(c, slots));
        std::vector<double> s(nba);
        PDX_HIP(c, hipMemcpy(s.data(), c->scf, sizeof(double) * nba, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < slots.size(); k++) scfactor[k] = s[slots[k]];
    }
    return PDX_OK;
}
