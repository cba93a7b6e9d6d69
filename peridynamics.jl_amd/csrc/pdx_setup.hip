// Context, body set-up and the bond system of libpdx.
//
// Bond system (reference: src/discretization/bond_system.jl:55-132, 142-176):
//  - points are reordered along a Morton curve (host, once) so that the neighbours of a
//    point are close in memory; every get/set maps back to the caller's order;
//  - the neighbour search runs on the device over a cell list with cell size = horizon
//    (the GridNeighborhoodSearch of bond_system.jl:97-99); one wavefront per point counts
//    the neighbours, a host prefix sum gives the CSR row pointers, and a second pass fills
//    the bonds of every point ordered by ascending caller index;
//  - the pair distance is formed as PointNeighbors does: pos_diff = x_i - x_j,
//    d2 = (p0*p0 + p1*p1) + p2*p2, accepted if d2 <= delta*delta, L = sqrt(d2), so that L
//    is bit-identical to the reference's Bond.length;
//  - precracks (predefined_cracks.jl:31-61, bond_system.jl:142-165, 249-267) are applied
//    inside the same passes: filtered bonds are never emitted, broken ones start inactive.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <thread>

#include "pdx_internal.h"

static const char *k_version = "pdx-mi355x 0.1.0 (gfx950)";

extern "C" const char *pdx_version(void) { return k_version; }

pdx_status pdx_fail(pdx_ctx *c, pdx_status st, const char *fmt, ...)
{
    if (c) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return st;
}

extern "C" const char *pdx_last_error(const pdx_ctx *c)
{
    return c ? c->err.c_str() : "null context";
}

#define FREE(p)                                                                             \
    do {                                                                                    \
        if (p) (void)hipFree(p);                                                            \
        p = nullptr;                                                                        \
    } while (0)

static void free_bonds(pdx_ctx *c)
{
    FREE(c->slice_ptr);
    FREE(c->cnt);
    FREE(c->bw);
    FREE(c->bw16);
    FREE(c->halo_ptr);
    FREE(c->halo);
    c->lds_mode = false;
    c->halo_max = 0;
    FREE(c->L);
    FREE(c->Lf);
}

void pdx_free_all(pdx_ctx *c)
{
    pdx_nk_free(c);
    pdx_gnn_free(c);
    FREE(c->d_psets);
    FREE(c->d_pid);
    FREE(c->scf);
    FREE(c->scf32);
    FREE(c->mfactor);
    FREE(c->fd_pts);
    FREE(c->fd_src);
    FREE(c->dr_vh_old);
    FREE(c->dr_bint_old);
    FREE(c->dr_density);
    FREE(c->dr_part);
    FREE(c->rec[0]);
    FREE(c->rec[1]);
    FREE(c->Xr);
    FREE(c->u);
    FREE(c->vh);
    FREE(c->v);
    FREE(c->a);
    FREE(c->bint);
    FREE(c->bext);
    FREE(c->damage);
    FREE(c->nact);
    FREE(c->killed);
    FREE(c->pstate);
    FREE(c->defgrad);
    FREE(c->cauchy);
    FREE(c->d_perm);
    FREE(c->d_iperm);
    FREE(c->stage);
    FREE(c->bcslot);
    FREE(c->slot_src);
    FREE(c->bcvals);
    free_bonds(c);
    c->dev_bytes = 0;
    c->bonds_built = false;
    c->bcvals_cap = 0;
}

extern "C" pdx_status pdx_create(pdx_ctx **out, int32_t device, int32_t precision)
{
    if (!out) return PDX_ERR_ARG;
    *out = nullptr;
    if (precision != PDX_FP64 && precision != PDX_MIXED) return PDX_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PDX_ERR_HIP;
    if (device < 0 || device >= ndev) return PDX_ERR_ARG;
    pdx_ctx *c = new pdx_ctx();
    c->device = device;
    c->precision = precision;
    if (const char *e = getenv("PDX_BB_UNROLL")) {
        int v = atoi(e);
        if (v == 2 || v == 4) c->bb_unroll = v;
    }
    if (const char *e = getenv("PDX_LDS_BUDGET")) c->lds_budget_env = atoi(e);
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return PDX_ERR_HIP;
    }
    if (hipMalloc((void **)&c->d_flags, 8 * sizeof(unsigned long long)) != hipSuccess) {
        delete c;
        return PDX_ERR_OOM;
    }
    *out = c;
    return PDX_OK;
}

extern "C" void pdx_destroy(pdx_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    pdx_export_free(c);
    pdx_free_all(c);
    FREE(c->d_flags);
    (void)pdx_timed_flush(c);
    for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
    if (c->open_ev) (void)hipEventDestroy(c->open_ev);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" pdx_status pdx_synchronize(pdx_ctx *c)
{
    if (!c) return PDX_ERR_ARG;
    PDX_HIP(c, hipSetDevice(c->device));
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    PDX_HIP(c, hipDeviceSynchronize());
    return PDX_OK;
}

extern "C" pdx_status pdx_set_points(pdx_ctx *c, int64_t n, const double *X, const double *vol,
                                     const uint8_t *fp)
{
    if (!c) return PDX_ERR_ARG;
    if (n <= 0 || !X || !vol) return pdx_fail(c, PDX_ERR_ARG, "invalid points (n=%lld)", (long long)n);
    if (n >= PDX_MAX_POINTS) return pdx_fail(c, PDX_ERR_UNSUPPORTED, "too many points");
    PDX_HIP(c, hipSetDevice(c->device));
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    pdx_free_all(c);
    c->n = n;
    c->nb = 0;
    c->X.assign(X, X + 3 * n);
    c->vol.assign(vol, vol + n);
    c->fail_permit.assign(n, 1);
    if (fp)
        for (int64_t i = 0; i < n; i++) c->fail_permit[i] = fp[i] ? 1 : 0;
    for (int64_t k = 0; k < 3 * n; k++)
        if (!std::isfinite(X[k])) return pdx_fail(c, PDX_ERR_ARG, "non-finite position");
    c->cracks.clear();
    c->n_src = 0;
    c->n_bcpts = 0;
    return PDX_OK;
}

extern "C" pdx_status pdx_set_material(pdx_ctx *c, int32_t kind, const pdx_params *p)
{
    if (!c || !p) return PDX_ERR_ARG;
    if (kind != PDX_BB && kind != PDX_OSB && kind != PDX_CPD)
        return pdx_fail(c, PDX_ERR_ARG, "unknown material kind %d", kind);
    if (!(p->delta > 0) || !(p->rho > 0))
        return pdx_fail(c, PDX_ERR_ARG, "horizon and rho have to be larger than zero");
    if (kind == PDX_CPD && p->zem != PDX_ZEM_SILLING && p->zem != PDX_ZEM_WAN)
        return pdx_fail(c, PDX_ERR_ARG, "unknown zero-energy-mode stabilisation %d", p->zem);
    if (kind == PDX_CPD && p->model != PDX_SVK && p->model != PDX_NHP && p->model != PDX_LINEAR_ELASTIC &&
        p->model != PDX_NEO_HOOKE)
        return pdx_fail(c, PDX_ERR_ARG, "unknown constitutive model %d", p->model);
    if (p->kernel < PDX_KERNEL_LINEAR || p->kernel > PDX_KERNEL_CUBIC_BSPLINE_NORM)
        return pdx_fail(c, PDX_ERR_ARG, "unknown influence function %d", p->kernel);
    if (kind == PDX_BB && p->kernel != PDX_KERNEL_LINEAR)
        return pdx_fail(c, PDX_ERR_ARG, "BBMaterial has no influence function");
    if (c->bonds_built && c->kind != kind)
        return pdx_fail(c, PDX_ERR_STATE, "material kind changed after the bonds were built");
    c->kind = kind;
    c->P = *p;
    return PDX_OK;
}

extern "C" pdx_status pdx_set_material_sets(pdx_ctx *c, int32_t kind, int32_t n_sets,
                                            const pdx_params *sets, const int32_t *point_set)
{
    if (!c || !sets || n_sets <= 0) return PDX_ERR_ARG;
    if (n_sets > PDX_MAX_PSETS) return pdx_fail(c, PDX_ERR_UNSUPPORTED, "at most %d parameter sets", PDX_MAX_PSETS);
    if (c->n <= 0) return pdx_fail(c, PDX_ERR_STATE, "pdx_set_points first");
    if (n_sets > 1 && !point_set) return PDX_ERR_ARG;
    PDX_TRY(pdx_set_material(c, kind, &sets[0]));
    c->psets.clear();
    c->pset_of.clear();
    if (n_sets == 1) return PDX_OK;
    if (kind == PDX_CPD)
        for (int32_t k = 1; k < n_sets; k++)
            if (sets[k].model != sets[0].model || sets[k].zem_cs != sets[0].zem_cs ||
                sets[k].maxdmg != sets[0].maxdmg || sets[k].zem != sets[0].zem)
                return pdx_fail(c, PDX_ERR_ARG, "the parameter sets of a CMaterial share its model, zem and maxdmg");
    for (int32_t k = 0; k < n_sets; k++) {
        if (!(sets[k].delta > 0) || !(sets[k].rho > 0))
            return pdx_fail(c, PDX_ERR_ARG, "horizon and rho have to be larger than zero");
        if (sets[k].delta != sets[0].delta)
            return pdx_fail(c, PDX_ERR_UNSUPPORTED, "the parameter sets of a body must share the horizon");
        if (sets[k].kernel != sets[0].kernel)
            return pdx_fail(c, PDX_ERR_ARG, "the parameter sets of a body share the material's influence function");
    }
    std::vector<uint8_t> of(c->n);
    for (int64_t i = 0; i < c->n; i++) {
        if (point_set[i] < 0 || point_set[i] >= n_sets)
            return pdx_fail(c, PDX_ERR_ARG, "parameter set of point %lld out of range", (long long)i);
        of[i] = (uint8_t)point_set[i];
    }
    c->psets.assign(sets, sets + n_sets);
    c->pset_of = of;
    return PDX_OK;
}

// device copies of the parameter sets and of each point's set (internal order)
static pdx_status upload_psets(pdx_ctx *c)
{
    if (!c->multi()) return PDX_OK;
    std::vector<pdx_pset> t(c->psets.size());
    for (size_t k = 0; k < t.size(); k++) {
        const pdx_params &q = c->psets[k];
        t[k] = {q.bc, q.epsc, q.epsc - 1e-11 * fabs(q.epsc), q.epsc + 1e-11 * fabs(q.epsc), q.rho,
                q.delta, q.G, q.K, q.lambda, q.mu, (q.zem_cs * q.bc) / q.delta, q.nu};
    }
    std::vector<uint8_t> pid(c->n);
    for (int64_t i = 0; i < c->n; i++) pid[c->iperm[i]] = c->pset_of[i];
    PDX_TRY(pdx_alloc(c, &c->d_psets, t.size()));
    PDX_TRY(pdx_alloc(c, &c->d_pid, c->n));
    PDX_HIP(c, hipMemcpy(c->d_psets, t.data(), sizeof(pdx_pset) * t.size(), hipMemcpyHostToDevice));
    PDX_HIP(c, hipMemcpy(c->d_pid, pid.data(), c->n, hipMemcpyHostToDevice));
    return PDX_OK;
}

extern "C" pdx_status pdx_add_precrack(pdx_ctx *c, int64_t na, const int32_t *a, int64_t nb,
                                       const int32_t *b, int32_t filter)
{
    if (!c) return PDX_ERR_ARG;
    if (c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "precracks must precede pdx_build_bonds");
    if ((int)c->cracks.size() >= PDX_MAX_CRACKS)
        return pdx_fail(c, PDX_ERR_UNSUPPORTED, "at most %d precracks", PDX_MAX_CRACKS);
    pdx_crack k;
    for (int64_t i = 0; i < na; i++) {
        if (a[i] < 0 || a[i] >= c->n) return pdx_fail(c, PDX_ERR_ARG, "precrack point out of range");
        k.a.push_back(a[i]);
    }
    for (int64_t i = 0; i < nb; i++) {
        if (b[i] < 0 || b[i] >= c->n) return pdx_fail(c, PDX_ERR_ARG, "precrack point out of range");
        k.b.push_back(b[i]);
    }
    k.filter = filter ? 1 : 0;
    c->cracks.push_back(k);
    return PDX_OK;
}

// ----------------------------------------------------------------------------------------
// Morton order (host, once per body)
// ----------------------------------------------------------------------------------------
static inline uint64_t spread3(uint64_t v)
{
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

static void morton_order(const std::vector<double> &X, int64_t n, std::vector<int32_t> &perm)
{
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            mn[d] = std::min(mn[d], X[3 * i + d]);
            mx[d] = std::max(mx[d], X[3 * i + d]);
        }
    double ext = std::max({mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2], 1e-300});
    double s = 2097151.0 / ext;
    std::vector<std::pair<uint64_t, int32_t>> key(n);
    for (int64_t i = 0; i < n; i++) {
        uint64_t q[3];
        for (int d = 0; d < 3; d++) {
            double t = (X[3 * i + d] - mn[d]) * s;
            q[d] = (uint64_t)std::min(std::max(t, 0.0), 2097151.0);
        }
        key[i] = {spread3(q[0]) | spread3(q[1]) << 1 | spread3(q[2]) << 2, (int32_t)i};
    }
    std::sort(key.begin(), key.end());
    perm.resize(n);
    for (int64_t m = 0; m < n; m++) perm[m] = key[m].second;
}

// ----------------------------------------------------------------------------------------
// neighbour search kernels
// ----------------------------------------------------------------------------------------
struct search_args {
    int64_t n;
    const double4 *Xr;        // internal order
    const int32_t *pcell;     // 3 per point
    const int64_t *cell_start;
    const int32_t *cell_pts;
    int64_t gx, gy, gz;
    double r2;                // delta * delta
    const int32_t *perm;      // caller index of each internal point
    const uint16_t *cmask;    // crack membership (bit 2k: set a of crack k, 2k+1: set b)
    uint32_t filt_bits;       // bit 2k set: crack k filters its bonds
    const uint8_t *fp;        // fail_permit, internal order
    int32_t *counts;
    const int64_t *slice_ptr;
    const int32_t *cnt;
    uint32_t *bw;
    double *L;
    unsigned long long *flags;
};

// (A_i & B_j) | (B_i & A_j) per crack, at the even bit positions
__device__ __forceinline__ uint32_t crack_hits(uint32_t mi, uint32_t mj)
{
    uint32_t ai = mi & 0x5555u, bi = (mi >> 1) & 0x5555u;
    uint32_t aj = mj & 0x5555u, bj = (mj >> 1) & 0x5555u;
    return (ai & bj) | (bi & aj);
}

__device__ __forceinline__ double pair_d2(double4 xi, double4 xj)
{
    double p0 = xi.x - xj.x, p1 = xi.y - xj.y, p2 = xi.z - xj.z;
    return (p0 * p0 + p1 * p1) + p2 * p2;
}

__global__ __launch_bounds__(PDX_BLOCK) void k_count_bonds(search_args s)
{
    int64_t i = pdx_wave_point();
    if (i >= s.n) return;
    int lane = threadIdx.x & 63;
    double4 xi = s.Xr[i];
    int cx = s.pcell[3 * i], cy = s.pcell[3 * i + 1], cz = s.pcell[3 * i + 2];
    uint32_t mi = s.cmask ? s.cmask[i] : 0;
    int cnt = 0;
    for (int z = cz - 1; z <= cz + 1; z++) {
        if (z < 0 || z >= s.gz) continue;
        for (int y = cy - 1; y <= cy + 1; y++) {
            if (y < 0 || y >= s.gy) continue;
            for (int x = cx - 1; x <= cx + 1; x++) {
                if (x < 0 || x >= s.gx) continue;
                int64_t cell = x + s.gx * (y + s.gy * (int64_t)z);
                int64_t e = s.cell_start[cell + 1];
                for (int64_t k = s.cell_start[cell] + lane; k < e; k += 64) {
                    int32_t j = s.cell_pts[k];
                    if (j == i) continue;
                    double d2 = pair_d2(xi, s.Xr[j]);
                    if (d2 <= s.r2) {
                        if (sqrt(d2) < 2.220446049250313e-16) {
                            if (atomicExch(&s.flags[1], 1ULL) == 0ULL) {
                                s.flags[2] = (unsigned long long)s.perm[i];
                                s.flags[3] = (unsigned long long)s.perm[j];
                            }
                        }
                        uint32_t h = s.cmask ? crack_hits(mi, s.cmask[j]) : 0u;
                        if (!(h & s.filt_bits)) cnt++;
                    }
                }
            }
        }
    }
    cnt = pdx_wave_sum_i(cnt);
    if (lane == 0) s.counts[i] = cnt;
}

__global__ __launch_bounds__(PDX_BLOCK) void k_fill_bonds(search_args s)
{
    __shared__ int32_t cj[PDX_WAVES_PER_BLOCK][PDX_FILL_CAP];
    __shared__ int32_t co[PDX_WAVES_PER_BLOCK][PDX_FILL_CAP];
    int64_t i = pdx_wave_point();
    if (i >= s.n) return;
    int lane = threadIdx.x & 63;
    int w = threadIdx.x >> 6;
    double4 xi = s.Xr[i];
    int cx = s.pcell[3 * i], cy = s.pcell[3 * i + 1], cz = s.pcell[3 * i + 2];
    uint32_t mi = s.cmask ? s.cmask[i] : 0;
    const int64_t b0 = s.slice_ptr[i >> 6] + (i & 63);  // SELL-64: bond t at b0 + 64 t
    const int want = s.cnt[i];                          // <= PDX_FILL_CAP, checked on the host
    int base = 0;
    for (int z = cz - 1; z <= cz + 1; z++) {
        if (z < 0 || z >= s.gz) continue;
        for (int y = cy - 1; y <= cy + 1; y++) {
            if (y < 0 || y >= s.gy) continue;
            for (int x = cx - 1; x <= cx + 1; x++) {
                if (x < 0 || x >= s.gx) continue;
                int64_t cell = x + s.gx * (y + s.gy * (int64_t)z);
                int64_t st = s.cell_start[cell], e = s.cell_start[cell + 1];
                for (int64_t k0 = st; k0 < e; k0 += 64) {
                    int64_t k = k0 + lane;
                    bool take = false;
                    int32_t j = -1;
                    if (k < e) {
                        j = s.cell_pts[k];
                        if (j != i) {
                            double d2 = pair_d2(xi, s.Xr[j]);
                            uint32_t h = s.cmask ? crack_hits(mi, s.cmask[j]) : 0u;
                            take = (d2 <= s.r2) && !(h & s.filt_bits);
                        }
                    }
                    uint64_t m = __ballot(take);
                    int pos = base + (int)__popcll(m & ((1ULL << lane) - 1ULL));
                    if (take && pos < PDX_FILL_CAP) {
                        cj[w][pos] = j;
                        co[w][pos] = s.perm[j];
                    }
                    base += (int)__popcll(m);
                }
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int cnt = min(base, want);
    for (int k = lane; k < cnt; k += 64) {
        // rows are ordered by the neighbour's internal (Morton) index, so that the lanes of
        // a gather read neighbouring records; pdx_get_bonds reorders rows by caller index
        int32_t ok = cj[w][k];
        int rank = 0;
        for (int m2 = 0; m2 < cnt; m2++) rank += (cj[w][m2] < ok) ? 1 : 0;
        int32_t j = cj[w][k];
        double4 xj = s.Xr[j];
        double d2 = pair_d2(xi, xj);
        uint32_t h = s.cmask ? crack_hits(mi, s.cmask[j]) : 0u;
        const uint32_t fpb = (s.fp[i] & s.fp[j]) ? PDX_BW_FP : 0u;
        const uint32_t act = (h & ~s.filt_bits) ? 0u : PDX_BW_ACT;  // broken at t = 0 by a precrack
        const int64_t b = b0 + (int64_t)rank * 64;
        s.bw[b] = (uint32_t)j | act | fpb;
        s.L[b] = sqrt(d2);
    }
}

__global__ void k_init_point_state(int64_t n, const int64_t *slice_ptr, const int32_t *cnt,
                                   const uint32_t *bw, int32_t *nact, double *damage)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t b0 = slice_ptr[i >> 6] + (i & 63);
    int na = 0;
    for (int t = 0; t < cnt[i]; t++) na += (bw[b0 + (int64_t)t * 64] & PDX_BW_ACT) ? 1 : 0;
    nact[i] = na;
    // calc_damage!: 1 - n_active / n_neighbors (NaN for a point without bonds)
    damage[i] = 1.0 - (double)na / (double)cnt[i];
}

__global__ void k_L_to_float(int64_t nb, const double *L, float *Lf)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nb) Lf[k] = (float)L[k];
}

// ----------------------------------------------------------------------------------------
// permuted transfers
// ----------------------------------------------------------------------------------------
// dev[m*stride + d] = stage[perm[m]*comps + d]
__global__ void k_scatter_in(int64_t n, int comps, int stride, const int32_t *perm,
                             const double *stage, double *dev)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * comps) return;
    int64_t m = t / comps;
    int d = (int)(t - m * comps);
    dev[m * stride + d] = stage[(int64_t)perm[m] * comps + d];
}

// stage[i*comps + d] = dev[iperm[i]*stride + d]
__global__ void k_gather_out(int64_t n, int comps, int stride, const int32_t *iperm,
                             const double *dev, double *stage)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * comps) return;
    int64_t i = t / comps;
    int d = (int)(t - i * comps);
    stage[t] = dev[(int64_t)iperm[i] * stride + d];
}

pdx_status pdx_upload_vec(pdx_ctx *c, const double *host, double *dev, int comps)
{
    int stride = comps;
    return pdx_upload_strided(c, host, dev, comps, stride);
}

pdx_status pdx_upload_strided(pdx_ctx *c, const double *host, double *dev, int comps, int stride)
{
    int64_t cnt = c->n * comps;
    PDX_HIP(c, hipMemcpyAsync(c->stage, host, sizeof(double) * cnt, hipMemcpyHostToDevice,
                              c->stream));
    int64_t blocks = (cnt + 255) / 256;
    hipLaunchKernelGGL(k_scatter_in, dim3(blocks), dim3(256), 0, c->stream, c->n, comps, stride,
                       c->d_perm, c->stage, dev);
    PDX_HIP(c, hipGetLastError());
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    return PDX_OK;
}

pdx_status pdx_download_vec(pdx_ctx *c, const double *dev, double *host, int comps, int stride)
{
    int64_t cnt = c->n * comps;
    int64_t blocks = (cnt + 255) / 256;
    hipLaunchKernelGGL(k_gather_out, dim3(blocks), dim3(256), 0, c->stream, c->n, comps, stride,
                       c->d_iperm, dev, c->stage);
    PDX_HIP(c, hipGetLastError());
    PDX_HIP(c, hipMemcpyAsync(host, c->stage, sizeof(double) * cnt, hipMemcpyDeviceToHost,
                              c->stream));
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    return PDX_OK;
}

// ----------------------------------------------------------------------------------------
// LDS halos of the BB step.  For every block of PDX_BLOCK consecutive (Morton) points the
// sorted, de-duplicated list of the points its bonds reach; every bond is rewritten as a
// 16-bit slot into its block's list.  Built on the host, in parallel over blocks, once.
// Returns false (and leaves the global-gather layout in use) if a halo does not fit.
// ----------------------------------------------------------------------------------------
static pdx_status build_halos(pdx_ctx *c, bool *ok)
{
    *ok = false;
    const int64_t n = c->n, nba = c->nb_alloc;
    const int64_t n_slices = (n + 63) / 64;
    const int64_t nblk = (n + PDX_BLOCK - 1) / PDX_BLOCK;
    hipDeviceProp_t prop;
    PDX_HIP(c, hipGetDeviceProperties(&prop, c->device));
    const int64_t lds_limit = std::min<int64_t>((int64_t)prop.sharedMemPerBlock, 160 * 1024);
    c->lds_limit = lds_limit - 64;
    // 64 bytes of slack for the alignment pad of a trailing byte table.  A table longer than
    // what LDS holds is still served: the kernels stage what fits and read the rest from
    // global memory (or stage it in tiles), so only the 13-bit slot field limits a table.
    const int64_t slot_limit = PDX_W16_SLOT + 1;
    std::vector<int64_t> sp(n_slices + 1);
    std::vector<int32_t> cn(n);
    std::vector<uint32_t> w(nba);
    PDX_HIP(c, hipMemcpy(sp.data(), c->slice_ptr, sizeof(int64_t) * (n_slices + 1), hipMemcpyDeviceToHost));
    PDX_HIP(c, hipMemcpy(cn.data(), c->cnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    PDX_HIP(c, hipMemcpy(w.data(), c->bw, sizeof(uint32_t) * nba, hipMemcpyDeviceToHost));
    std::vector<std::vector<int32_t>> tab(nblk);
    std::vector<uint16_t> w16(nba, 0);
    std::vector<int> too_big(nblk, 0), unsorted(nblk, 0);
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto work = [&](int64_t k0, int64_t k1) {
        std::vector<int32_t> tmp;
        for (int64_t k = k0; k < k1; k++) {
            tmp.clear();
            const int64_t pe = std::min(n, (k + 1) * PDX_BLOCK);
            for (int64_t p = k * PDX_BLOCK; p < pe; p++) {
                const int64_t b0 = sp[p >> 6] + (p & 63);
                for (int t = 0; t < cn[p]; t++) tmp.push_back((int32_t)(w[b0 + 64 * (int64_t)t] & PDX_BW_NBR));
            }
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
            if ((int64_t)tmp.size() > slot_limit) {
                too_big[k] = 1;
                continue;
            }
            for (int64_t p = k * PDX_BLOCK; p < pe; p++) {
                const int64_t b0 = sp[p >> 6] + (p & 63);
                int64_t prev = -1;
                for (int t = 0; t < cn[p]; t++) {
                    const int64_t b = b0 + 64 * (int64_t)t;
                    const int32_t j = (int32_t)(w[b] & PDX_BW_NBR);
                    const uint32_t slot = (uint32_t)(std::lower_bound(tmp.begin(), tmp.end(), j) - tmp.begin());
                    if ((int64_t)slot <= prev) unsorted[k] = 1;
                    prev = slot;
                    w16[b] = (uint16_t)(slot | ((w[b] & PDX_BW_ACT) ? PDX_W16_ACT | PDX_W16_RACT : 0u) |
                                        ((w[b] & PDX_BW_FP) ? PDX_W16_FP : 0u));
                }
            }
            tab[k] = tmp;
        }
    };
    std::vector<std::thread> th;
    const int64_t per = (nblk + nth - 1) / nth;
    for (unsigned t = 0; t < nth; t++) {
        const int64_t k0 = t * per, k1 = std::min(nblk, k0 + per);
        if (k0 < k1) th.emplace_back(work, k0, k1);
    }
    for (auto &t : th) t.join();
    // keep the global-gather kernels if a table outgrows the slot field, or if a row's slots
    // do not increase (the tiled staging walks every row in slot order)
    for (int64_t k = 0; k < nblk; k++)
        if (too_big[k] || unsorted[k]) return PDX_OK;
    std::vector<int64_t> hp(nblk + 1, 0);
    int64_t hmax = 0;
    for (int64_t k = 0; k < nblk; k++) {
        hp[k + 1] = hp[k] + (int64_t)tab[k].size();
        hmax = std::max<int64_t>(hmax, (int64_t)tab[k].size());
    }
    std::vector<int32_t> halo(std::max<int64_t>(hp[nblk], 1));
    for (int64_t k = 0; k < nblk; k++) std::copy(tab[k].begin(), tab[k].end(), halo.begin() + hp[k]);
    PDX_TRY(pdx_alloc(c, &c->bw16, nba));
    PDX_TRY(pdx_alloc(c, &c->halo_ptr, nblk + 1));
    PDX_TRY(pdx_alloc(c, &c->halo, halo.size()));
    PDX_HIP(c, hipMemcpy(c->bw16, w16.data(), sizeof(uint16_t) * nba, hipMemcpyHostToDevice));
    PDX_HIP(c, hipMemcpy(c->halo_ptr, hp.data(), sizeof(int64_t) * (nblk + 1), hipMemcpyHostToDevice));
    PDX_HIP(c, hipMemcpy(c->halo, halo.data(), sizeof(int32_t) * halo.size(), hipMemcpyHostToDevice));
    c->halo_max = (int32_t)hmax;
    c->halo_total = hp[nblk];
    c->lds_mode = true;
    *ok = true;
    return PDX_OK;
}

// ----------------------------------------------------------------------------------------
// pdx_build_bonds
// ----------------------------------------------------------------------------------------
extern "C" pdx_status pdx_build_bonds(pdx_ctx *c, int64_t *n_bonds_out)
{
    if (!c) return PDX_ERR_ARG;
    if (c->n <= 0) return pdx_fail(c, PDX_ERR_STATE, "pdx_set_points first");
    if (c->kind < 0) return pdx_fail(c, PDX_ERR_STATE, "pdx_set_material first");
    PDX_HIP(c, hipSetDevice(c->device));
    const int64_t n = c->n;
    if (c->bonds_built) {
        // rebuilding: drop everything but the points
        std::vector<pdx_crack> keep = c->cracks;
        pdx_free_all(c);
        c->cracks = keep;
    }
    // Morton order
    morton_order(c->X, n, c->perm);
    c->iperm.assign(n, 0);
    for (int64_t m = 0; m < n; m++) c->iperm[c->perm[m]] = (int32_t)m;
    std::vector<double4> Xm(n);
    std::vector<uint8_t> fpm(n);
    for (int64_t m = 0; m < n; m++) {
        int64_t i = c->perm[m];
        Xm[m] = make_double4(c->X[3 * i], c->X[3 * i + 1], c->X[3 * i + 2], c->vol[i]);
        fpm[m] = c->fail_permit[i];
    }
    // cell list over the internal order, cell size = horizon
    double cs = c->P.delta;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t m = 0; m < n; m++) {
        const double v3[3] = {Xm[m].x, Xm[m].y, Xm[m].z};
        for (int d = 0; d < 3; d++) {
            mn[d] = std::min(mn[d], v3[d]);
            mx[d] = std::max(mx[d], v3[d]);
        }
    }
    for (;;) {
        double ncell = 1;
        for (int d = 0; d < 3; d++) ncell *= std::floor((mx[d] - mn[d]) / cs) + 1.0;
        if (ncell <= 8.0 * (double)n + 4096.0) break;
        cs *= 2.0;  // a coarser grid still finds every neighbour within the horizon
    }
    int64_t g[3];
    for (int d = 0; d < 3; d++) g[d] = (int64_t)std::floor((mx[d] - mn[d]) / cs) + 1;
    int64_t ncell = g[0] * g[1] * g[2];
    std::vector<int32_t> pcell(3 * n);
    std::vector<int64_t> cell_start(ncell + 1, 0);
    std::vector<int64_t> cid(n);
    for (int64_t m = 0; m < n; m++) {
        const double v3[3] = {Xm[m].x, Xm[m].y, Xm[m].z};
        for (int d = 0; d < 3; d++) {
            int64_t q = (int64_t)std::floor((v3[d] - mn[d]) / cs);
            q = std::min(std::max(q, (int64_t)0), g[d] - 1);
            pcell[3 * m + d] = (int32_t)q;
        }
        cid[m] = pcell[3 * m] + g[0] * (pcell[3 * m + 1] + g[1] * (int64_t)pcell[3 * m + 2]);
        cell_start[cid[m] + 1]++;
    }
    for (int64_t k = 0; k < ncell; k++) cell_start[k + 1] += cell_start[k];
    std::vector<int32_t> cell_pts(n);
    {
        std::vector<int64_t> fill(cell_start.begin(), cell_start.end() - 1);
        for (int64_t m = 0; m < n; m++) cell_pts[fill[cid[m]]++] = (int32_t)m;
    }
    // crack membership masks
    std::vector<uint16_t> cmask;
    uint32_t filt_bits = 0;
    if (!c->cracks.empty()) {
        cmask.assign(n, 0);
        for (size_t k = 0; k < c->cracks.size(); k++) {
            for (int32_t i : c->cracks[k].a) cmask[c->iperm[i]] |= (uint16_t)(1u << (2 * k));
            for (int32_t i : c->cracks[k].b) cmask[c->iperm[i]] |= (uint16_t)(1u << (2 * k + 1));
            if (c->cracks[k].filter) filt_bits |= 1u << (2 * k);
        }
    }
    // device buffers
    PDX_TRY(pdx_alloc(c, &c->rec[0], n));
    PDX_TRY(pdx_alloc(c, &c->rec[1], n));
    PDX_TRY(pdx_alloc(c, &c->Xr, n));
    PDX_TRY(pdx_alloc(c, &c->d_perm, n));
    PDX_TRY(pdx_alloc(c, &c->d_iperm, n));
    PDX_TRY(pdx_alloc(c, &c->stage, 9 * n));
    for (double **p : {&c->u, &c->vh, &c->v, &c->a, &c->bint, &c->bext})
        PDX_TRY(pdx_alloc(c, p, 3 * n));
    PDX_TRY(pdx_alloc(c, &c->damage, n));
    PDX_TRY(pdx_alloc(c, &c->nact, n));
    PDX_TRY(pdx_alloc(c, &c->killed, n));
    PDX_TRY(pdx_alloc(c, &c->bcslot, n));
    int32_t *d_pcell = nullptr, *d_cell_pts = nullptr, *d_counts = nullptr;
    int64_t *d_cell_start = nullptr;
    uint16_t *d_cmask = nullptr;
    uint8_t *d_fp = nullptr;
    auto cleanup = [&]() {
        FREE(d_pcell);
        FREE(d_cell_pts);
        FREE(d_counts);
        FREE(d_cell_start);
        FREE(d_cmask);
        FREE(d_fp);
    };
#define TRY_C(expr)                                                                         \
    do {                                                                                    \
        pdx_status s_ = (expr);                                                             \
        if (s_ != PDX_OK) {                                                                 \
            cleanup();                                                                      \
            return s_;                                                                      \
        }                                                                                   \
    } while (0)
#define HIP_C(call)                                                                         \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            cleanup();                                                                      \
            return pdx_fail(c, PDX_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
        }                                                                                   \
    } while (0)
    HIP_C(hipMalloc((void **)&d_pcell, sizeof(int32_t) * 3 * n));
    HIP_C(hipMalloc((void **)&d_cell_pts, sizeof(int32_t) * n));
    HIP_C(hipMalloc((void **)&d_counts, sizeof(int32_t) * n));
    HIP_C(hipMalloc((void **)&d_cell_start, sizeof(int64_t) * (ncell + 1)));
    HIP_C(hipMalloc((void **)&d_fp, n));
    if (!cmask.empty()) HIP_C(hipMalloc((void **)&d_cmask, sizeof(uint16_t) * n));
    HIP_C(hipMemcpy(c->Xr, Xm.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
    HIP_C(hipMemcpy(c->rec[0], Xm.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
    HIP_C(hipMemcpy(c->rec[1], Xm.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
    HIP_C(hipMemcpy(c->d_perm, c->perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    HIP_C(hipMemcpy(c->d_iperm, c->iperm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    HIP_C(hipMemcpy(d_pcell, pcell.data(), sizeof(int32_t) * 3 * n, hipMemcpyHostToDevice));
    HIP_C(hipMemcpy(d_cell_pts, cell_pts.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    HIP_C(hipMemcpy(d_cell_start, cell_start.data(), sizeof(int64_t) * (ncell + 1),
                    hipMemcpyHostToDevice));
    HIP_C(hipMemcpy(d_fp, fpm.data(), n, hipMemcpyHostToDevice));
    if (d_cmask)
        HIP_C(hipMemcpy(d_cmask, cmask.data(), sizeof(uint16_t) * n, hipMemcpyHostToDevice));
    HIP_C(hipMemset(c->d_flags, 0, 8 * sizeof(unsigned long long)));
    // Each (Xr) record keeps the volume in .w for the bond passes; the reference record's
    // .w holds the volume too (only .xyz are read as the reference position).

    search_args s{};
    s.n = n;
    s.Xr = c->Xr;
    s.pcell = d_pcell;
    s.cell_start = d_cell_start;
    s.cell_pts = d_cell_pts;
    s.gx = g[0];
    s.gy = g[1];
    s.gz = g[2];
    s.r2 = c->P.delta * c->P.delta;
    s.perm = c->d_perm;
    s.cmask = d_cmask;
    s.filt_bits = filt_bits;
    s.fp = d_fp;
    s.counts = d_counts;
    s.flags = c->d_flags;
    int64_t grid = pdx_wave_grid(n);
    hipLaunchKernelGGL(k_count_bonds, dim3(grid), dim3(PDX_BLOCK), 0, c->stream, s);
    HIP_C(hipGetLastError());
    std::vector<int32_t> counts(n);
    HIP_C(hipMemcpyAsync(counts.data(), d_counts, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                         c->stream));
    unsigned long long hflags[8];
    HIP_C(hipMemcpyAsync(hflags, c->d_flags, sizeof hflags, hipMemcpyDeviceToHost, c->stream));
    HIP_C(hipStreamSynchronize(c->stream));
    if (hflags[1]) {
        cleanup();
        return pdx_fail(c, PDX_ERR_DUPLICATE_POINT,
                        "point duplicate found! Point #%llu has a duplicate #%llu which will lead to NaNs",
                        hflags[2] + 1, hflags[3] + 1);
    }
    // SELL-64 slices: slice s holds points 64 s .. 64 s + 63, as wide as its widest row
    const int64_t n_slices = (n + 63) / 64;
    std::vector<int64_t> slice_ptr(n_slices + 1, 0);
    int32_t maxc = 0;
    int64_t nb_real = 0;
    for (int64_t sl = 0; sl < n_slices; sl++) {
        int32_t w = 0;
        for (int64_t m = 64 * sl; m < std::min(n, 64 * sl + 64); m++) {
            w = std::max(w, counts[m]);
            nb_real += counts[m];
        }
        maxc = std::max(maxc, w);
        slice_ptr[sl + 1] = slice_ptr[sl] + 64 * (int64_t)w;
    }
    if (maxc > PDX_FILL_CAP) {
        cleanup();
        return pdx_fail(c, PDX_ERR_UNSUPPORTED, "a point has %d bonds (> %d)", maxc, PDX_FILL_CAP);
    }
    const int64_t nb = nb_real;
    const int64_t nb_alloc = slice_ptr[n_slices];
    if (nb_alloc >= ((int64_t)1 << 40)) {
        cleanup();
        return pdx_fail(c, PDX_ERR_UNSUPPORTED, "too many bonds");
    }
    c->nb = nb;
    c->nb_alloc = nb_alloc;
    TRY_C(pdx_alloc(c, &c->slice_ptr, n_slices + 1));
    TRY_C(pdx_alloc(c, &c->cnt, n));
    TRY_C(pdx_alloc(c, &c->bw, nb_alloc));
    TRY_C(pdx_alloc(c, &c->L, nb_alloc));
    HIP_C(hipMemcpy(c->slice_ptr, slice_ptr.data(), sizeof(int64_t) * (n_slices + 1),
                    hipMemcpyHostToDevice));
    HIP_C(hipMemcpy(c->cnt, counts.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    // padding entries of a slice stay zero (never read: t < cnt)
    HIP_C(hipMemsetAsync(c->bw, 0, sizeof(uint32_t) * nb_alloc, c->stream));
    HIP_C(hipMemsetAsync(c->L, 0, sizeof(double) * nb_alloc, c->stream));
    s.slice_ptr = c->slice_ptr;
    s.cnt = c->cnt;
    s.bw = c->bw;
    s.L = c->L;
    hipLaunchKernelGGL(k_fill_bonds, dim3(grid), dim3(PDX_BLOCK), 0, c->stream, s);
    HIP_C(hipGetLastError());
    hipLaunchKernelGGL(k_init_point_state, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       c->stream, n, c->slice_ptr, c->cnt, c->bw, c->nact, c->damage);
    HIP_C(hipGetLastError());
    // zero dynamic state
    for (double *p : {c->u, c->vh, c->v, c->a, c->bint, c->bext})
        HIP_C(hipMemsetAsync(p, 0, sizeof(double) * 3 * n, c->stream));
    HIP_C(hipMemsetAsync(c->killed, 0, n, c->stream));
    HIP_C(hipMemsetAsync(c->bcslot, 0xff, sizeof(int32_t) * n, c->stream));
    HIP_C(hipStreamSynchronize(c->stream));
    c->bext_nonzero = false;
    c->cur = 0;
    c->n_src = 0;
    c->n_bcpts = 0;
    // material state between the passes
    // OSB: the per-point pair of pass 1, then (LDS layout, one set) the own force part
    if (c->kind == PDX_OSB) TRY_C(pdx_alloc(c, &c->pstate, 5 * n));
    if (c->kind == PDX_CPD) {
        // Q per point; with ZEMWan also the stabilisation stiffness C_1 (15 numbers per point).
        // The LDS passes keep six more per point behind those: the own bond terms and
        // a = Q X + beta x (ZEMWan: Q X + C_1 x)
        TRY_C(pdx_alloc(c, &c->pstate, (c->P.zem == PDX_ZEM_WAN ? 21 : 15) * n));
        TRY_C(pdx_alloc(c, &c->defgrad, 9 * n));
        TRY_C(pdx_alloc(c, &c->cauchy, 9 * n));
        HIP_C(hipMemsetAsync(c->defgrad, 0, sizeof(double) * 9 * n, c->stream));
        HIP_C(hipMemsetAsync(c->cauchy, 0, sizeof(double) * 9 * n, c->stream));
        HIP_C(hipStreamSynchronize(c->stream));
    }
    cleanup();
    // LDS-staged layouts: BB (x|V per neighbour-table entry), OSB (x|V and c1, c2), CPD
    // (tables staged in tiles)
    if (!getenv("PDX_NO_LDS")) {
        bool ok = false;
        PDX_TRY(build_halos(c, &ok));
    }
    // FP32 lengths for the global-gather mixed BB step (the LDS-staged steps take L = |D|)
    if (c->precision == PDX_MIXED && !c->lds_mode) {
        const int64_t nba = c->nb_alloc;
        PDX_TRY(pdx_alloc(c, &c->Lf, nba));
        hipLaunchKernelGGL(k_L_to_float, dim3((unsigned)((nba + 255) / 256)), dim3(256), 0, c->stream,
                           nba, c->L, c->Lf);
        PDX_HIP(c, hipGetLastError());
    }
    PDX_TRY(upload_psets(c));
    if (c->kind == PDX_CPD && c->P.zem == PDX_ZEM_WAN && !c->lds_mode)
        return pdx_fail(c, PDX_ERR_UNSUPPORTED,
                        "ZEMWan needs the LDS-staged bond layout (a neighbour table did not fit)");
    if (c->multi() && !c->lds_mode)
        return pdx_fail(c, PDX_ERR_UNSUPPORTED,
                        "several parameter sets need the LDS-staged bond layout (a neighbour table did not fit)");
    if (c->corr && !c->lds_mode)
        return pdx_fail(c, PDX_ERR_UNSUPPORTED,
                        "the surface correction needs the LDS-staged bond layout (BB or OSB, neighbour tables that fit)");
    PDX_TRY(pdx_build_correction(c));
    c->bonds_built = true;
    if (n_bonds_out) *n_bonds_out = nb;
    return PDX_OK;
#undef TRY_C
#undef HIP_C
}

pdx_status pdx_caller_bond_slots(pdx_ctx *c, std::vector<int64_t> &slots)
{
    const int64_t n = c->n, nba = c->nb_alloc;
    const int64_t n_slices = (n + 63) / 64;
    std::vector<int64_t> sp(n_slices + 1);
    std::vector<int32_t> cn(n);
    std::vector<uint32_t> w(nba);
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    PDX_HIP(c, hipMemcpy(sp.data(), c->slice_ptr, sizeof(int64_t) * (n_slices + 1), hipMemcpyDeviceToHost));
    PDX_HIP(c, hipMemcpy(cn.data(), c->cnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    PDX_HIP(c, hipMemcpy(w.data(), c->bw, sizeof(uint32_t) * nba, hipMemcpyDeviceToHost));
    slots.clear();
    std::vector<std::pair<int32_t, int64_t>> row;
    for (int64_t i = 0; i < n; i++) {
        const int64_t m = c->iperm[i];
        const int64_t b0 = sp[m >> 6] + (m & 63);
        row.clear();
        for (int t = 0; t < cn[m]; t++) {
            const int64_t b = b0 + (int64_t)t * 64;
            row.push_back({c->perm[w[b] & PDX_BW_NBR], b});
        }
        std::sort(row.begin(), row.end());
        for (auto &e : row) slots.push_back(e.second);
    }
    return PDX_OK;
}

extern "C" pdx_status pdx_get_bonds(pdx_ctx *c, int64_t *row_ptr, int32_t *nbr, double *L,
                                    uint8_t *fp, uint8_t *active)
{
    if (!c) return PDX_ERR_ARG;
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    const int64_t n = c->n, nba = c->nb_alloc;
    const int64_t n_slices = (n + 63) / 64;
    std::vector<int64_t> sp(n_slices + 1);
    std::vector<int32_t> cn(n);
    std::vector<uint32_t> w(nba);
    std::vector<double> L_dev(L ? nba : 0);
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    PDX_HIP(c, hipMemcpy(sp.data(), c->slice_ptr, sizeof(int64_t) * (n_slices + 1),
                         hipMemcpyDeviceToHost));
    PDX_HIP(c, hipMemcpy(cn.data(), c->cnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    PDX_HIP(c, hipMemcpy(w.data(), c->bw, sizeof(uint32_t) * nba, hipMemcpyDeviceToHost));
    if (L) PDX_HIP(c, hipMemcpy(L_dev.data(), c->L, sizeof(double) * nba, hipMemcpyDeviceToHost));
    if (c->lds_mode) {
        // the LDS-staged step keeps the bond state in the 16-bit words
        std::vector<uint16_t> w16(nba);
        PDX_HIP(c, hipMemcpy(w16.data(), c->bw16, sizeof(uint16_t) * nba, hipMemcpyDeviceToHost));
        for (int64_t b = 0; b < nba; b++)
            w[b] = (w[b] & ~PDX_BW_ACT) | ((w16[b] & PDX_W16_ACT) ? PDX_BW_ACT : 0u);
    }
    std::vector<uint8_t> killed(n, 0);
    if (c->kind == PDX_CPD)
        PDX_HIP(c, hipMemcpy(killed.data(), c->killed, n, hipMemcpyDeviceToHost));
    // caller order
    std::vector<int64_t> out_rp(n + 1, 0);
    for (int64_t i = 0; i < n; i++) {
        int64_t m = c->iperm[i];
        out_rp[i + 1] = out_rp[i] + cn[m];
    }
    if (row_ptr) std::memcpy(row_ptr, out_rp.data(), sizeof(int64_t) * (n + 1));
    // device rows are in Morton order of the neighbours; the caller's rows list neighbours
    // by ascending caller index (the reference's bond order within a grid cell)
    std::vector<std::pair<int32_t, int64_t>> row;
    for (int64_t i = 0; i < n; i++) {
        int64_t m = c->iperm[i];
        int64_t o = out_rp[i];
        row.clear();
        const int64_t b0 = sp[m >> 6] + (m & 63);
        for (int t = 0; t < cn[m]; t++) {
            const int64_t b = b0 + (int64_t)t * 64;
            row.push_back({c->perm[w[b] & PDX_BW_NBR], b});
        }
        std::sort(row.begin(), row.end());
        for (auto &e : row) {
            const int64_t b = e.second;
            if (nbr) nbr[o] = e.first;
            if (L) L[o] = L_dev[b];
            if (fp) fp[o] = (w[b] & PDX_BW_FP) ? 1 : 0;
            if (active) active[o] = ((w[b] & PDX_BW_ACT) && !killed[m]) ? 1 : 0;
            o++;
        }
    }
    return PDX_OK;
}

// ----------------------------------------------------------------------------------------
// fields
// ----------------------------------------------------------------------------------------
pdx_status pdx_field_ptr(pdx_ctx *c, int field, double **p, int *comps, int *stride)
{
    *p = nullptr;
    *comps = 3;
    *stride = 3;
    switch (field) {
    case PDX_FIELD_POSITION: *p = (double *)c->rec[c->cur]; *stride = 4; break;
    case PDX_FIELD_DISPLACEMENT: *p = c->u; break;
    case PDX_FIELD_VELOCITY: *p = c->v; break;
    case PDX_FIELD_VELOCITY_HALF: *p = c->vh; break;
    case PDX_FIELD_ACCELERATION: *p = c->a; break;
    case PDX_FIELD_B_INT: *p = c->bint; break;
    case PDX_FIELD_B_EXT: *p = c->bext; break;
    case PDX_FIELD_DAMAGE: *p = c->damage; *comps = 1; *stride = 1; break;
    case PDX_FIELD_DEFGRAD: *p = c->defgrad; *comps = 9; *stride = 9; break;
    case PDX_FIELD_CAUCHY_STRESS: *p = c->cauchy; *comps = 9; *stride = 9; break;
    case PDX_FIELD_REF_POSITION: *p = (double *)c->Xr; *stride = 4; break;
    case PDX_FIELD_VOLUME: *p = ((double *)c->Xr) + 3; *comps = 1; *stride = 4; break;
    case PDX_FIELD_VELOCITY_HALF_OLD: *p = c->dr_vh_old; break;
    case PDX_FIELD_B_INT_OLD: *p = c->dr_bint_old; break;
    case PDX_FIELD_DENSITY_MATRIX: *p = c->dr_density; *comps = 1; *stride = 1; break;
    default: return pdx_fail(c, PDX_ERR_ARG, "unknown field %d", field);
    }
    if (!*p) return pdx_fail(c, PDX_ERR_ARG, "field %d is not held by this material", field);
    return PDX_OK;
}

__global__ void k_int_to_double(int64_t n, const int32_t *src, double *dst)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = (double)src[k];
}

__global__ void k_copy_xyz(int64_t n, const double4 *src, double4 *dst)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k];
}

pdx_status pdx_field_comps(pdx_ctx *c, int field, int *comps)
{
    if (field == PDX_FIELD_N_ACTIVE_BONDS) {
        *comps = 1;
        return PDX_OK;
    }
    double *p;
    int stride;
    PDX_TRY(pdx_field_ptr(c, field, &p, comps, &stride));
    return PDX_OK;
}

pdx_status pdx_gather_field(pdx_ctx *c, int field, double *dst)
{
    double *p;
    int comps = 1, stride = 1;
    if (field == PDX_FIELD_N_ACTIVE_BONDS) {
        double *tmp = c->stage + 4 * c->n;  // second half of the staging buffer
        hipLaunchKernelGGL(k_int_to_double, dim3((c->n + 255) / 256), dim3(256), 0, c->stream,
                           c->n, c->nact, tmp);
        hipLaunchKernelGGL(k_gather_out, dim3((c->n + 255) / 256), dim3(256), 0, c->stream,
                           c->n, 1, 1, c->d_iperm, tmp, dst);
        PDX_HIP(c, hipGetLastError());
        return PDX_OK;
    }
    PDX_TRY(pdx_field_ptr(c, field, &p, &comps, &stride));
    const int64_t cnt = c->n * comps;
    hipLaunchKernelGGL(k_gather_out, dim3((cnt + 255) / 256), dim3(256), 0, c->stream, c->n, comps,
                       stride, c->d_iperm, p, dst);
    PDX_HIP(c, hipGetLastError());
    return PDX_OK;
}

extern "C" pdx_status pdx_get_field(pdx_ctx *c, int32_t field, double *host)
{
    if (!c || !host) return PDX_ERR_ARG;
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    PDX_HIP(c, hipSetDevice(c->device));
    if (field == PDX_FIELD_N_ACTIVE_BONDS) {
        double *tmp = c->stage + 4 * c->n;  // second half of the staging buffer
        hipLaunchKernelGGL(k_int_to_double, dim3((c->n + 255) / 256), dim3(256), 0, c->stream,
                           c->n, c->nact, tmp);
        PDX_HIP(c, hipGetLastError());
        return pdx_download_vec(c, tmp, host, 1, 1);
    }
    double *p;
    int comps, stride;
    PDX_TRY(pdx_field_ptr(c, field, &p, &comps, &stride));
    if (field == PDX_FIELD_VOLUME) {
        // .w of the reference record
        return pdx_download_vec(c, p, host, 1, 4);
    }
    return pdx_download_vec(c, p, host, comps, stride);
}

extern "C" pdx_status pdx_set_field(pdx_ctx *c, int32_t field, const double *host)
{
    if (!c || !host) return PDX_ERR_ARG;
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    if (field == PDX_FIELD_REF_POSITION || field == PDX_FIELD_VOLUME ||
        field == PDX_FIELD_N_ACTIVE_BONDS || field == PDX_FIELD_DEFGRAD ||
        field == PDX_FIELD_CAUCHY_STRESS || field == PDX_FIELD_DENSITY_MATRIX)
        return pdx_fail(c, PDX_ERR_ARG, "field %d is read only", field);
    PDX_HIP(c, hipSetDevice(c->device));
    double *p;
    int comps, stride;
    PDX_TRY(pdx_field_ptr(c, field, &p, &comps, &stride));
    PDX_TRY(pdx_upload_strided(c, host, p, comps, stride));
    if (field == PDX_FIELD_B_EXT) {
        c->bext_nonzero = false;
        for (int64_t k = 0; k < 3 * c->n; k++)
            if (host[k] != 0.0) c->bext_nonzero = true;
    }
    return PDX_OK;
}

// ----------------------------------------------------------------------------------------
// stable time step: calc_timestep_point (bond_system.jl:269-277).  The device rows are in
// Morton order of the neighbours; the terms are ranked by the neighbour's caller index in
// LDS and summed in that order by one lane, so that the value equals the reference's bit
// for bit.
// ----------------------------------------------------------------------------------------
__global__ __launch_bounds__(PDX_BLOCK) void k_timestep(int64_t n, const int64_t *slice_ptr,
                                                         const int32_t *cnts,
                                                         const uint32_t *bw, const double *L,
                                                         const double4 *Xr, const int32_t *perm,
                                                         double bc, double rho,
                                                         const pdx_pset *psets, const uint8_t *pid,
                                                         double *dt)
{
    __shared__ double term[PDX_WAVES_PER_BLOCK][PDX_FILL_CAP];
    __shared__ int32_t key[PDX_WAVES_PER_BLOCK][PDX_FILL_CAP];
    const int64_t i = pdx_wave_point();
    if (i >= n) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t b0 = slice_ptr[i >> 6] + (i & 63);
    const int cnt = cnts[i];  // <= PDX_FILL_CAP (checked when built)
    if (psets) {  // the point's own parameter set (get_params(chunk, point_id))
        bc = psets[pid[i]].bc;
        rho = psets[pid[i]].rho;
    }
    for (int k = lane; k < cnt; k += 64) {
        const uint32_t j = bw[b0 + (int64_t)k * 64] & PDX_BW_NBR;
        key[w][k] = perm[j];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int k = lane; k < cnt; k += 64) {
        const int64_t b = b0 + (int64_t)k * 64;
        const uint32_t j = bw[b] & PDX_BW_NBR;
        const int32_t kk = key[w][k];
        int rank = 0;
        for (int m = 0; m < cnt; m++) rank += (key[w][m] < kk) ? 1 : 0;
        term[w][rank] = Xr[j].w * bc / L[b];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
        double dtsum = 0.0;
        for (int k = 0; k < cnt; k++) dtsum += term[w][k];
        dt[i] = sqrt(2.0 * rho / dtsum);
    }
}

extern "C" pdx_status pdx_stable_timestep(pdx_ctx *c, double safety, double *dt_out)
{
    if (!c || !dt_out) return PDX_ERR_ARG;
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    if (!(safety > 0 && safety < 1))
        return pdx_fail(c, PDX_ERR_ARG, "`safety_factor` of VelocityVerlet has invalid value!");
    PDX_HIP(c, hipSetDevice(c->device));
    double *tmp = c->stage;
    hipLaunchKernelGGL(k_timestep, dim3((unsigned)pdx_wave_grid(c->n)), dim3(PDX_BLOCK), 0,
                       c->stream, c->n, c->slice_ptr, c->cnt, c->bw, c->L, c->Xr, c->d_perm, c->P.bc,
                       c->P.rho, c->multi() ? c->d_psets : nullptr, c->multi() ? c->d_pid : nullptr, tmp);
    PDX_HIP(c, hipGetLastError());
    std::vector<double> dt(c->n);
    PDX_HIP(c, hipMemcpyAsync(dt.data(), tmp, sizeof(double) * c->n, hipMemcpyDeviceToHost,
                              c->stream));
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    double mn = INFINITY;
    for (double d : dt) mn = std::min(mn, d);
    *dt_out = mn * safety;
    return PDX_OK;
}

// ----------------------------------------------------------------------------------------
// velocity boundary conditions
// ----------------------------------------------------------------------------------------
extern "C" pdx_status pdx_set_velocity_bcs(pdx_ctx *c, int64_t npts, const int32_t *points,
                                           const int32_t *src, int64_t n_src)
{
    if (!c) return PDX_ERR_ARG;
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    if (npts < 0 || n_src < 0 || (npts > 0 && (!points || !src)))
        return pdx_fail(c, PDX_ERR_ARG, "invalid boundary condition table");
    PDX_HIP(c, hipSetDevice(c->device));
    std::vector<int32_t> slot(c->n, -1);
    std::vector<int32_t> ssrc(3 * std::max<int64_t>(npts, 1), -1);
    for (int64_t k = 0; k < npts; k++) {
        int32_t i = points[k];
        if (i < 0 || i >= c->n) return pdx_fail(c, PDX_ERR_ARG, "BC point out of range");
        if (slot[c->iperm[i]] >= 0) return pdx_fail(c, PDX_ERR_ARG, "BC point listed twice");
        slot[c->iperm[i]] = (int32_t)k;
        for (int d = 0; d < 3; d++) {
            int32_t sidx = src[3 * k + d];
            if (sidx < -1 || sidx >= n_src) return pdx_fail(c, PDX_ERR_ARG, "BC source out of range");
            ssrc[3 * k + d] = sidx;
        }
    }
    if (c->slot_src) {
        (void)hipFree(c->slot_src);
        c->slot_src = nullptr;
    }
    PDX_TRY(pdx_alloc(c, &c->slot_src, ssrc.size()));
    PDX_HIP(c, hipMemcpy(c->bcslot, slot.data(), sizeof(int32_t) * c->n, hipMemcpyHostToDevice));
    PDX_HIP(c, hipMemcpy(c->slot_src, ssrc.data(), sizeof(int32_t) * ssrc.size(),
                         hipMemcpyHostToDevice));
    c->n_bcpts = npts;
    c->n_src = n_src;
    return PDX_OK;
}

// force-density conditions on b_ext (forcedensity_bc! with f(t) or f(p, t),
// boundary_conditions.jl:559-575): dof 3 points[k] + d takes value src[3k + d] of the per-step
// row (the same rows as the velocity conditions, whose length pdx_set_velocity_bcs set)
extern "C" pdx_status pdx_set_force_density_bcs(pdx_ctx *c, int64_t npts, const int32_t *points,
                                                const int32_t *src)
{
    if (!c) return PDX_ERR_ARG;
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    if (npts < 0 || (npts > 0 && (!points || !src)))
        return pdx_fail(c, PDX_ERR_ARG, "invalid boundary condition table");
    PDX_HIP(c, hipSetDevice(c->device));
    std::vector<int32_t> pts(std::max<int64_t>(npts, 1), 0), ssrc(3 * std::max<int64_t>(npts, 1), -1);
    std::vector<uint8_t> seen(c->n, 0);
    for (int64_t k = 0; k < npts; k++) {
        const int32_t i = points[k];
        if (i < 0 || i >= c->n) return pdx_fail(c, PDX_ERR_ARG, "BC point out of range");
        if (seen[i]) return pdx_fail(c, PDX_ERR_ARG, "BC point listed twice");
        seen[i] = 1;
        pts[k] = c->iperm[i];
        for (int d = 0; d < 3; d++) {
            const int32_t sidx = src[3 * k + d];
            if (sidx < -1 || sidx >= c->n_src) return pdx_fail(c, PDX_ERR_ARG, "BC source out of range");
            ssrc[3 * k + d] = sidx;
        }
    }
    if (c->fd_pts) (void)hipFree(c->fd_pts);
    if (c->fd_src) (void)hipFree(c->fd_src);
    c->fd_pts = nullptr;
    c->fd_src = nullptr;
    PDX_TRY(pdx_alloc(c, &c->fd_pts, pts.size()));
    PDX_TRY(pdx_alloc(c, &c->fd_src, ssrc.size()));
    PDX_HIP(c, hipMemcpy(c->fd_pts, pts.data(), sizeof(int32_t) * pts.size(), hipMemcpyHostToDevice));
    PDX_HIP(c, hipMemcpy(c->fd_src, ssrc.data(), sizeof(int32_t) * ssrc.size(), hipMemcpyHostToDevice));
    c->n_fdpts = npts;
    if (npts > 0) c->bext_nonzero = true;
    return PDX_OK;
}

// b_ext of the constrained dofs from one row of condition values (non-finite: unchanged)
__global__ void k_fd_bcs(int64_t npts, const int32_t *pts, const int32_t *src, const double *bcv,
                         double *bext)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * npts) return;
    const int64_t k = t / 3;
    const int d = (int)(t - 3 * k);
    const int32_t s = src[t];
    if (s < 0) return;
    const double val = bcv[s];
    if (isfinite(val)) bext[3 * (int64_t)pts[k] + d] = val;
}

pdx_status pdx_apply_fd_bcs(pdx_ctx *c, const double *bcv_row)
{
    if (c->n_fdpts <= 0 || !bcv_row) return PDX_OK;
    hipLaunchKernelGGL(k_fd_bcs, dim3((unsigned)((3 * c->n_fdpts + 255) / 256)), dim3(256), 0, c->stream,
                       c->n_fdpts, c->fd_pts, c->fd_src, bcv_row, c->bext);
    PDX_HIP(c, hipGetLastError());
    return PDX_OK;
}

// ----------------------------------------------------------------------------------------
// instrumentation
// ----------------------------------------------------------------------------------------
extern "C" pdx_status pdx_set_profiling(pdx_ctx *c, int32_t enable)
{
    if (!c) return PDX_ERR_ARG;
    PDX_TRY(pdx_timed_flush(c));
    c->prof = enable != 0;
    c->ktime.clear();
    return PDX_OK;
}

static pdx_status take_event(pdx_ctx *c, hipEvent_t *e)
{
    if (!c->evpool.empty()) {
        *e = c->evpool.back();
        c->evpool.pop_back();
        return PDX_OK;
    }
    PDX_HIP(c, hipEventCreate(e));
    return PDX_OK;
}

// Launch timing with HIP events on the context's stream.  Events are only recorded here;
// they are read (and the stream waited for) in pdx_timed_flush, so timing does not
// serialise back-to-back launches.
pdx_status pdx_timed_begin(pdx_ctx *c)
{
    if (!c->prof) return PDX_OK;
    PDX_TRY(take_event(c, &c->open_ev));
    PDX_HIP(c, hipEventRecord(c->open_ev, c->stream));
    return PDX_OK;
}

pdx_status pdx_timed_end(pdx_ctx *c, const char *name)
{
    if (!c->prof || !c->open_ev) return PDX_OK;
    hipEvent_t b;
    PDX_TRY(take_event(c, &b));
    PDX_HIP(c, hipEventRecord(b, c->stream));
    c->pending.push_back({name, c->open_ev, b});
    c->open_ev = nullptr;
    if (c->pending.size() >= 4096) PDX_TRY(pdx_timed_flush(c));
    return PDX_OK;
}

pdx_status pdx_timed_flush(pdx_ctx *c)
{
    if (c->pending.empty()) return PDX_OK;
    PDX_HIP(c, hipEventSynchronize(c->pending.back().b));
    for (auto &t : c->pending) {
        float ms = 0;
        PDX_HIP(c, hipEventElapsedTime(&ms, t.a, t.b));
        auto &e = c->ktime[t.name];
        e.first += ms;
        e.second += 1;
        c->evpool.push_back(t.a);
        c->evpool.push_back(t.b);
    }
    c->pending.clear();
    return PDX_OK;
}

extern "C" pdx_status pdx_kernel_time(pdx_ctx *c, const char *name, double *avg_ms, int64_t *count)
{
    if (!c || !name) return PDX_ERR_ARG;
    PDX_TRY(pdx_timed_flush(c));
    auto it = c->ktime.find(name);
    if (it == c->ktime.end()) {
        if (avg_ms) *avg_ms = 0;
        if (count) *count = 0;
        return PDX_OK;
    }
    if (avg_ms) *avg_ms = it->second.second ? it->second.first / it->second.second : 0.0;
    if (count) *count = it->second.second;
    return PDX_OK;
}

extern "C" pdx_status pdx_info(pdx_ctx *c, int64_t *np, int64_t *nb, int64_t *bytes)
{
    if (!c) return PDX_ERR_ARG;
    if (np) *np = c->n;
    if (nb) *nb = c->nb;
    if (bytes) *bytes = c->dev_bytes;
    return PDX_OK;
}

extern "C" pdx_status pdx_layout(pdx_ctx *c, int32_t *lds_mode, int64_t *bond_slots,
                                 int64_t *halo_entries, int32_t *halo_max)
{
    if (!c) return PDX_ERR_ARG;
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    if (lds_mode) *lds_mode = c->lds_mode ? 1 : 0;
    if (bond_slots) *bond_slots = c->nb_alloc;
    if (halo_entries) *halo_entries = c->lds_mode ? c->halo_total : 0;
    if (halo_max) *halo_max = c->lds_mode ? c->halo_max : 0;
    return PDX_OK;
}

pdx_status pdx_reset_flags(pdx_ctx *c)
{
    unsigned long long init[2] = {~0ULL, 0ULL};
    PDX_HIP(c, hipMemcpyAsync(c->d_flags, init, sizeof init, hipMemcpyHostToDevice, c->stream));
    return PDX_OK;
}

pdx_status pdx_check_flags(pdx_ctx *c, int64_t *nan_step)
{
    unsigned long long h[2];
    PDX_HIP(c, hipMemcpyAsync(h, c->d_flags, sizeof h, hipMemcpyDeviceToHost, c->stream));
    PDX_HIP(c, hipStreamSynchronize(c->stream));
    if (nan_step) *nan_step = (h[0] == ~0ULL) ? -1 : (int64_t)h[0];
    return PDX_OK;
}
