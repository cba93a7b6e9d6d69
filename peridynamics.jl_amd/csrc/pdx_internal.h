// Internal declarations of libpdx: the context, device helpers and kernel launchers.
// All device code targets gfx950 (CDNA4, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "pdx.h"

#define PDX_WAVE 64
#define PDX_BLOCK 256                      // 4 waves per block
#define PDX_WAVES_PER_BLOCK (PDX_BLOCK / PDX_WAVE)
#define PDX_MAX_CRACKS 8
#define PDX_FILL_CAP 1024                  // max bonds of one point in the search kernels
#define PDX_NUM_XCD 8

// bond word: bits 0-29 neighbour (internal index), bit 30 active, bit 31 fail_permit
#define PDX_BW_NBR 0x3fffffffu
#define PDX_BW_ACT 0x40000000u
#define PDX_BW_FP 0x80000000u
#define PDX_MAX_POINTS ((int64_t)1 << 30)
#define PDX_W16_SLOT 0x1fffu  // halo slot: <= 8192 entries (160 KiB LDS / 32 B per entry = 5120)
#define PDX_W16_RACT 0x2000u  // the reverse bond (neighbour -> owner) is active
#define PDX_W16_ACT 0x4000u
#define PDX_W16_FP 0x8000u

// flags of the step kernels
#define PDX_F_HAS_NEXT 1    // integrate to the next step (write positions of step n+1)
#define PDX_F_STORE 2       // materialise v, a, b_int, damage, n_active of this step
#define PDX_F_BEXT 4        // b_ext is not all zero
#define PDX_F_FORCE_ONLY 8  // calc_force_density! only: write b_int, damage, n_active
#define PDX_F_SCALE_VOL 16  // force-only: write b_int * V (Newton-Krylov residual form)
#define PDX_F_FP64 32       // FP64 kernels also in a mixed-precision context (Newton-Krylov)

struct pdx_crack {
    std::vector<int32_t> a, b;
    int filter;
};

struct pdx_nk;   // Newton-Krylov workspace (pdx_nk.hip)

// one parameter set on the device (bodies with several material! sets)
struct pdx_pset {
    double bc, epsc, epsc_lo, epsc_hi, rho, delta, G, K, lam, mu;
    double beta;  // ZEMSilling beta = Cs bc / delta of the set (correspondence.jl:232)
    double nu;    // Poisson's ratio (LinearElastic's Hooke matrix)
};
#define PDX_MAX_PSETS 255
// feature mask of the LDS-staged step kernels
#define PDX_FEAT_MULTI 1  // several parameter sets
#define PDX_FEAT_ZW 8     // CPD with ZEMWan (stabilisation stiffness per point)
#define PDX_FEAT_SC 2     // EnergySurfaceCorrection scfactor per bond
#define PDX_FEAT_CAP 4    // neighbour tables longer than the LDS budget: the rest from global memory
#define PDX_FEAT_KERN 16  // OSB with an influence function other than linear_kernel
struct pdx_gnn;  // learned preconditioner (pdx_gnn.hip)

struct pdx_ctx {
    int device = 0;
    int precision = PDX_FP64;
    // bonds per iteration of the BB thread (tuning knob PDX_BB_UNROLL = 2 | 4)
    int bb_unroll = 0;  // bonds per lane and unrolled iteration of the BB step: 2, 4, 0 = default
    int64_t lds_limit = 160 * 1024;  // dynamic LDS bytes a workgroup may use
    int32_t lds_budget_env = 0;      // PDX_LDS_BUDGET (bytes), 0 = default
    hipStream_t stream = nullptr;
    std::string err;

    int64_t n = 0;   // points
    int64_t nb = 0;  // bonds

    // host copies in the caller's order
    std::vector<double> X, vol;
    std::vector<uint8_t> fail_permit;
    std::vector<int32_t> perm;   // perm[m] = caller index of internal point m (Morton order)
    std::vector<int32_t> iperm;  // iperm[caller index] = internal index

    int kind = -1;
    pdx_params P{};
    // several parameter sets: the sets, each caller point's set, their device copies
    // (internal order); empty psets = the single set P
    std::vector<pdx_params> psets;
    std::vector<uint8_t> pset_of;
    pdx_pset *d_psets = nullptr;
    uint8_t *d_pid = nullptr;
    bool multi() const { return psets.size() > 1; }
    // EnergySurfaceCorrection (bond_system_corrections.jl:46-204): 0 none, 1 energy based
    int32_t corr = 0;
    double *scf = nullptr;     // scfactor per bond slot (SELL-64, like L)
    float *scf32 = nullptr;    // FP32 copy for the mixed step
    double *mfactor = nullptr; // 3 x n (internal order), mfactor[d * n + i]
    // export pipeline: two snapshot slots (device, pinned host) filled in turn; a snapshot
    // is gathered on the compute stream and copied to the host on xstream
    hipStream_t xstream = nullptr;
    hipEvent_t xev_snap[2] = {nullptr, nullptr}, xev_done[2] = {nullptr, nullptr};
    double *xdev[2] = {nullptr, nullptr}, *xhost[2] = {nullptr, nullptr};
    size_t xcap[2] = {0, 0};
    std::vector<int32_t> xfields[2];
    std::vector<int64_t> xoff[2];
    int64_t xtag[2] = {0, 0};
    int64_t xbegun = 0, xwaited = 0;
    std::vector<pdx_crack> cracks;
    bool bonds_built = false;

    // device state, internal (Morton) order
    double4 *rec[2] = {nullptr, nullptr};  // (x, y, z, volume) current positions, ping-pong
    int cur = 0;
    double4 *Xr = nullptr;                 // (X, Y, Z, 0) reference positions
    double *u = nullptr, *vh = nullptr, *v = nullptr, *a = nullptr, *bint = nullptr,
           *bext = nullptr;                // 3n each
    double *damage = nullptr;              // n
    int32_t *nact = nullptr;               // n
    uint8_t *killed = nullptr;             // n (CPD too_much_damage!)
    double *pstate = nullptr;              // per-point state between the two passes
    double *defgrad = nullptr, *cauchy = nullptr;  // 9n (CPD)
    int32_t *d_perm = nullptr, *d_iperm = nullptr;
    double *stage = nullptr;               // 9n staging for permuted transfers
    // bonds in sliced ELL (SELL-64): bond t of point p at slice_ptr[p / 64] + 64 t + p % 64
    int64_t *slice_ptr = nullptr;          // n_slices + 1
    int32_t *cnt = nullptr;                // n: bonds of each point
    int64_t nb_alloc = 0;                  // entries including the padding of the slices
    uint32_t *bw = nullptr;                // bond words: neighbour | active | fail_permit
    double *L = nullptr;                   // reference lengths (FP64 layout)
    float *Lf = nullptr;                   // reference lengths (mixed layout)
    // LDS-staged BB step: per block of PDX_BLOCK points, the sorted list of every point
    // its bonds reach (the halo); bonds then address the halo with 16-bit words
    // (bits 0-13 slot, bit 14 active, bit 15 fail_permit)
    bool lds_mode = false;
    uint16_t *bw16 = nullptr;              // nb_alloc
    int64_t *halo_ptr = nullptr;           // n_blocks + 1
    int32_t *halo = nullptr;               // halo entries (internal point indices)
    int32_t halo_max = 0;                  // largest halo of a block (LDS layout: every row's
                                           // table slots increase)
    int64_t halo_total = 0;                // sum of the halo sizes of all blocks
    bool bext_nonzero = false;

    // velocity boundary conditions
    int32_t *bcslot = nullptr;             // n, -1 = none
    int32_t *slot_src = nullptr;           // 3 per slot
    double *bcvals = nullptr;              // value table for a run
    int64_t n_src = 0, n_bcpts = 0, bcvals_cap = 0;
    int32_t *fd_pts = nullptr, *fd_src = nullptr;  // force-density conditions (internal points)
    int64_t n_fdpts = 0;

    unsigned long long *d_flags = nullptr; // [0] first NaN step, [1] duplicate flag, [2..3] pair

    // instrumentation
    bool prof = false;
    std::map<std::string, std::pair<double, int64_t>> ktime;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    struct timed_launch {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<timed_launch> pending;  // recorded, not yet read
    std::vector<hipEvent_t> evpool;     // idle events
    hipEvent_t open_ev = nullptr;       // begin event of the launch being timed

    pdx_nk *nk = nullptr;
    // dynamic relaxation (pdx_dr.hip)
    double *dr_vh_old = nullptr, *dr_bint_old = nullptr;  // 3n
    double *dr_density = nullptr;                           // n (Lambda, the same in x, y, z)
    double *dr_part = nullptr;                              // damping partial sums, cn
    bool dr_ready = false;
    pdx_gnn *gnn = nullptr;
    int64_t dev_bytes = 0;
};

// error helpers ---------------------------------------------------------------------------
pdx_status pdx_fail(pdx_ctx *c, pdx_status st, const char *fmt, ...);
#define PDX_HIP(c, call)                                                                    \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return pdx_fail((c), PDX_ERR_HIP, "%s failed: %s (%s:%d)", #call,               \
                            hipGetErrorString(e_), __FILE__, __LINE__);                     \
    } while (0)
#define PDX_TRY(expr)                                                                       \
    do {                                                                                    \
        pdx_status s_ = (expr);                                                             \
        if (s_ != PDX_OK) return s_;                                                        \
    } while (0)

template <typename T>
pdx_status pdx_alloc(pdx_ctx *c, T **p, size_t count)
{
    if (*p) return PDX_OK;
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return pdx_fail(c, PDX_ERR_OOM, "hipMalloc of %zu bytes failed: %s", bytes,
                        hipGetErrorString(e));
    }
    c->dev_bytes += (int64_t)bytes;
    return PDX_OK;
}
void pdx_free_all(pdx_ctx *c);

// launch helpers --------------------------------------------------------------------------
inline int64_t pdx_wave_grid(int64_t n_waves)
{
    int64_t blocks = (n_waves + PDX_WAVES_PER_BLOCK - 1) / PDX_WAVES_PER_BLOCK;
    return (blocks + PDX_NUM_XCD - 1) / PDX_NUM_XCD * PDX_NUM_XCD;  // multiple of 8
}

pdx_status pdx_timed_begin(pdx_ctx *c);
pdx_status pdx_timed_end(pdx_ctx *c, const char *name);
pdx_status pdx_timed_flush(pdx_ctx *c);

// device helpers --------------------------------------------------------------------------
// Influence function omega(delta, L) of the OSB and CPD bonds (kernels.jl), rL = 1/L.  The
// default linear_kernel delta / L is formed as delta * (1/L) (within an ulp); the others as
// Julia evaluates them: xi = L / delta, xi^2 = xi xi, xi^3 = xi xi xi, left to right.
__device__ __noinline__ double pdx_omega_other(int kid, double delta, double L)
{
    if (kid == PDX_KERNEL_CONST_ONE) return 1.0;  // const_one_kernel (:10)
    const double xi = L / delta;
    const double x2 = xi * xi, x3 = xi * xi * xi;
    if (kid == PDX_KERNEL_CUBIC_BSPLINE) {  // cubic_b_spline_kernel (:43-52)
        if (0.0 < xi && xi <= 0.5) return ((2.0 / 3.0) - 4.0 * x2) + 4.0 * x3;
        if (0.5 < xi && xi <= 1.0) return (((4.0 / 3.0) - 4.0 * xi) + 4.0 * x2) - (4.0 / 3.0) * x3;
        return 0.0;
    }
    // cubic_b_spline_kernel_norm (:74-84)
    const double C = 8.0 / (3.141592653589793 * (delta * delta * delta));
    if (0.0 < xi && xi <= 0.5) return C * ((1.0 - 6.0 * x2) + 6.0 * x3);
    if (0.5 < xi && xi <= 1.0) {
        const double t = 1.0 - xi;
        return (C * 2.0) * (t * t * t);
    }
    return 0.0;
}

__device__ __forceinline__ double pdx_omega(int kid, double delta, double L, double rL)
{
    return kid == PDX_KERNEL_LINEAR ? delta * rL : pdx_omega_other(kid, delta, L);
}

// finite-difference perturbation of J v (compute_perturbation, newton_krylov.jl:534-566):
// eps = scale cbrt(eps_mach) (||u|| + 1) / ||v||, with sc = scale cbrt(eps_mach) and the
// squared norms on the device (a vanishing ||v|| counts as 1)
__device__ __forceinline__ double pdx_nk_perturbation_v(double un2, double vn2, double sc)
{
    double vn = sqrt(vn2);
    if (vn < 2.220446049250313e-16) vn = 1.0;
    return sc * (sqrt(un2) + 1.0) / vn;
}
__device__ __forceinline__ double pdx_nk_perturbation(const double *un2, const double *vn2, double sc)
{
    return pdx_nk_perturbation_v(*un2, *vn2, sc);
}

// 32 sums per lane reduced over the wave by a halving butterfly: each step trades half of
// the values with the partner lane (32 shuffles in all instead of 6 per value); afterwards
// lanes 2q and 2q + 1 hold sum q in s[0].  Indices are compile-time constants, so the array
// stays in registers.
template <int H>
__device__ __forceinline__ void pdx_bfly32(double (&s)[32], int lane)
{
    const bool up = (lane & (2 * H)) != 0;
#pragma unroll
    for (int q = 0; q < H; q++) {
        const double keep = up ? s[q + H] : s[q];
        const double give = up ? s[q] : s[q + H];
        s[q] = keep + __shfl_xor(give, 2 * H);
    }
}

__device__ __forceinline__ void pdx_wave_reduce32(double (&s)[32])
{
    const int lane = threadIdx.x & 63;
    pdx_bfly32<16>(s, lane);
    pdx_bfly32<8>(s, lane);
    pdx_bfly32<4>(s, lane);
    pdx_bfly32<2>(s, lane);
    pdx_bfly32<1>(s, lane);
    s[0] += __shfl_xor(s[0], 1);
}

// sum of parts[0, np) by a whole workgroup: every thread returns it, and the order of the
// additions depends only on np and the workgroup size (sh: one double per wave)
__device__ __forceinline__ double pdx_block_sum_parts(const double *parts, int np, double *sh)
{
    double s = 0.0;
    for (int t = threadIdx.x; t < np; t += blockDim.x) s += parts[t];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    const int nw = (int)((blockDim.x + 63) >> 6);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    double tot = 0.0;
    for (int q = 0; q < nw; q++) tot += sh[q];
    __syncthreads();
    return tot;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.  Give each
// XCD one contiguous range of points so that its L2 holds only its own neighbourhoods.
// gridDim.x is a multiple of 8, so the map is a bijection of [0, gridDim.x).
__device__ __forceinline__ int64_t pdx_xcd_block()
{
    int64_t g = gridDim.x, b = blockIdx.x;
    int64_t per = g / PDX_NUM_XCD;
    return (b % PDX_NUM_XCD) * per + b / PDX_NUM_XCD;
}

// wave-uniform point index of the calling wave (one wave per point)
__device__ __forceinline__ int64_t pdx_wave_point()
{
    int64_t w = pdx_xcd_block() * PDX_WAVES_PER_BLOCK + (threadIdx.x >> 6);
    return __builtin_amdgcn_readfirstlane((int)w);
}

__device__ __forceinline__ double pdx_wave_sum(double x)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

__device__ __forceinline__ int pdx_wave_sum_i(int x)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

__device__ __forceinline__ float pdx_wave_sum_f(float x)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// DPP move of a double (two 32-bit halves); CTRL is a DPP control word
template <int CTRL>
__device__ __forceinline__ double pdx_dpp(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

template <int CTRL>
__device__ __forceinline__ int pdx_dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}

// sum over each row of 16 lanes, left in every lane of the row:
// quad_perm(1,0,3,2), quad_perm(2,3,0,1), row_ror:4, row_ror:8
__device__ __forceinline__ double pdx_row_sum(double x)
{
    x += pdx_dpp<0xB1>(x);
    x += pdx_dpp<0x4E>(x);
    x += pdx_dpp<0x124>(x);
    x += pdx_dpp<0x128>(x);
    return x;
}

__device__ __forceinline__ int pdx_row_sum_i(int x)
{
    x += pdx_dpp_i<0xB1>(x);
    x += pdx_dpp_i<0x4E>(x);
    x += pdx_dpp_i<0x124>(x);
    x += pdx_dpp_i<0x128>(x);
    return x;
}

__device__ __forceinline__ double pdx_readlane(double x, int l)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
    return __hiloint2double(hi, lo);
}

// sum over each half wave (lanes 0-31, 32-63), returned to the lanes of that half
__device__ __forceinline__ double pdx_half_sum(double x, int half)
{
    x = pdx_row_sum(x);
    const double r0 = pdx_readlane(x, 0), r1 = pdx_readlane(x, 16);
    const double r2 = pdx_readlane(x, 32), r3 = pdx_readlane(x, 48);
    return half ? (r2 + r3) : (r0 + r1);
}

__device__ __forceinline__ int pdx_half_sum_i(int x, int half)
{
    x = pdx_row_sum_i(x);
    const int r0 = __builtin_amdgcn_readlane(x, 0), r1 = __builtin_amdgcn_readlane(x, 16);
    const int r2 = __builtin_amdgcn_readlane(x, 32), r3 = __builtin_amdgcn_readlane(x, 48);
    return half ? (r2 + r3) : (r0 + r1);
}

// arguments of the per-step kernels (shared by BB, OSB and CPD)
struct pdx_step_args {
    int64_t n;
    int64_t step;
    int flags;
    double rho, dt, dth;
    const int64_t *slice_ptr;
    const int32_t *cnt;
    uint32_t *bw;
    uint16_t *bw16;
    const int64_t *halo_ptr;
    const int32_t *halo;
    int32_t lds_budget;  // LDS bytes a block stages its neighbour table into
    int32_t halo_max;
    const double *scf;  // EnergySurfaceCorrection scfactor per bond slot (FP64), or null
    const float *scf32; // the same in FP32 (mixed precision)
    const pdx_pset *psets;  // several parameter sets: the table and each point's set
    const uint8_t *pid;
    const double *L;
    const float *Lf;
    const double4 *rec_in;
    double4 *rec_out;
    const double4 *Xr;
    double *u, *vh;
    const double *bext;
    double *v_out, *a_out, *bint_out, *dmg_out;
    int32_t *nact_out;
    const int32_t *bcslot, *slot_src;
    const double *bcv;
    unsigned long long *flags_out;
    // material
    double bc, epsc, epsc_lo, epsc_hi, delta, G, K, lambda, mu, beta_zem, maxdmg, nu;
    int model;
    int kernel;  // influence function (PDX_KERNEL_*)
    double *pstate;
    uint8_t *killed;
    double *defgrad, *cauchy;
};

pdx_status pdx_fill_step_args(pdx_ctx *c, pdx_step_args *a, int64_t step, int flags, double dt,
                              const double *bcv_dev);
pdx_status pdx_launch_step(pdx_ctx *c, const pdx_step_args &a);  // one step of the material
// J v = (F(u + e v) V - F(u - e v) V) / (2 e) of the bond-based material in one kernel, the
// constrained rows (cmask) = v; true when the context's layout allows it (FP64 LDS step
// with one parameter set, no surface correction, a neighbour table that fits the LDS)
// ||v||^2 = the sum of the np partial sums `parts` (pdx_block_sum_parts at 1024 threads)
bool pdx_bb_jv_fusable(const pdx_ctx *c);
pdx_status pdx_launch_bb_jv(pdx_ctx *c, const double *u, const double *v, const double *un2,
                            const double *parts, int np, double sc, const uint8_t *cmask,
                            double *jv);
pdx_status pdx_launch_prologue(pdx_ctx *c, int64_t step, double dt, const double *bcv_dev);
pdx_status pdx_apply_fd_bcs(pdx_ctx *c, const double *bcv_row);  // b_ext conditions of one step
pdx_status pdx_build_correction(pdx_ctx *c);  // EnergySurfaceCorrection factors
// LDS bytes for a block's neighbour table: half the CU's LDS (two 1024-thread blocks per
// CU), minus room for the parameter-set bytes' alignment
inline int pdx_lds_budget(const pdx_ctx *c)
{
    const int64_t half = 80 * 1024 - 64;
    const int64_t b = c->lds_budget_env > 0 ? c->lds_budget_env : half;
    return (int)std::min<int64_t>(b, c->lds_limit);
}
// for kernels whose registers allow one 1024-thread block per CU anyway (the CPD passes):
// the whole LDS of the CU
inline int pdx_lds_budget_full(const pdx_ctx *c)
{
    const int64_t b = c->lds_budget_env > 0 ? c->lds_budget_env : c->lds_limit;
    return (int)std::min<int64_t>(b, c->lds_limit);
}
void pdx_export_free(pdx_ctx *c);              // the export pipeline's buffers and stream
// device pointer, components and stride of a field (PDX_FIELD_*)
pdx_status pdx_field_ptr(pdx_ctx *c, int field, double **p, int *comps, int *stride);
// components of a field, and its gather in the caller's point order into dst (n * comps
// doubles of device memory), enqueued on the compute stream
pdx_status pdx_field_comps(pdx_ctx *c, int field, int *comps);
pdx_status pdx_gather_field(pdx_ctx *c, int field, double *dst);
// device slot of every bond in the caller's bond order (rows by caller index, neighbours by
// ascending caller index)
pdx_status pdx_caller_bond_slots(pdx_ctx *c, std::vector<int64_t> &slots);
pdx_status pdx_check_flags(pdx_ctx *c, int64_t *nan_step);
pdx_status pdx_reset_flags(pdx_ctx *c);

// permuted transfers (pdx_setup.hip)
pdx_status pdx_upload_vec(pdx_ctx *c, const double *host, double *dev, int comps);
pdx_status pdx_upload_strided(pdx_ctx *c, const double *host, double *dev, int comps, int stride);
pdx_status pdx_download_vec(pdx_ctx *c, const double *dev, double *host, int comps, int stride);

// Newton-Krylov (pdx_nk.hip)
void pdx_nk_free(pdx_ctx *c);
void pdx_gnn_free(pdx_ctx *c);
