// Force density and the fused Velocity Verlet step: BB, OSB and CPD.
//
// Bond storage is sliced ELL with slices of 64 Morton-consecutive points (SELL-64): bond t
// of point p sits at slice_ptr[p / 64] + 64 t + p % 64, so bond t of the 64 points of a
// slice is one contiguous 64-entry row.  One thread owns one point: it walks its bonds
// with fully coalesced loads of the bond words and lengths, gathers the neighbours'
// records (Morton order keeps them in the L1/L2 of the CU), and accumulates its own
// force density in registers -- no atomics and no cross-lane reductions.
//
// The scatter of the reference's OSB and CPD force densities (b_int[j] -= ...,
// ordinary_state_based.jl:145-146, correspondence.jl:248-249) is rewritten as a gather: a
// first pass stores each point's state (OSB: c1, c2; CPD: F and P*K^-1), the second pass
// adds, for every bond i-j, both the term point i scatters into itself and the term point
// j scatters into i.  Failure of a bond depends only on the symmetric stretch, so the bond
// state of i-j and j-i agree; CPD's too_much_damage! (correspondence.jl:288-298) is kept
// as a per-point flag that masks the bonds of the killed point only.
//
// The Velocity Verlet step (velocity_verlet.jl:181-197) is fused into the force launch:
// launch n computes the force at x^n, then a^n and v^n (_update_acc_and_vel!, :262-295),
// then the half step of step n+1 (update_vel_half!, :239-245), its boundary conditions
// (apply_boundary_conditions!, boundary_conditions.jl:212-226) and u^{n+1}, x^{n+1}
// (update_disp_and_pos!, :247-255) into the other position buffer.  A prologue launch does
// the first half step, so a run of N steps is N + 1 launches.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <mutex>
#include <unordered_map>

#include "pdx_internal.h"

#define PDX_LDS_LANES 4  // lanes per point in the LDS-staged BB kernels

// thread index of the calling point; blocks are remapped so that each XCD gets a
// contiguous range of points (see pdx_xcd_block)
__device__ __forceinline__ int64_t pdx_thread_point()
{
    return pdx_xcd_block() * PDX_BLOCK + threadIdx.x;
}

// first bond entry of point p in the SELL-64 arrays; bond t is at + 64 t
// Neighbour-table entries a block stages in LDS: all of them up to a budget that leaves
// room for two 1024-thread blocks per CU; a block whose table is longer reads the entries
// past the budget from global memory (through the caches).  ENTRY = LDS bytes per entry.
__host__ __device__ __forceinline__ int pdx_lds_entries(int halo_max, int budget, int entry)
{
    const int cap = budget / entry;
    return halo_max < cap ? (halo_max > 0 ? halo_max : 1) : cap;
}

__device__ __forceinline__ int64_t pdx_bond0(const int64_t *slice_ptr, int64_t p)
{
    return slice_ptr[p >> 6] + (p & 63);
}

// Bond words and lengths of G bonds of a lane's row: t, t + S, ..., t + (G - 1) S.  Bonds
// past the row's end read its last bond (bond 0 of an empty row), an in-range address whose
// value the caller never uses, so the loads need no branch and can be issued a whole group
// ahead of their use.
template <int G, int S>
__device__ __forceinline__ void pdx_load_group(const uint16_t *wbase, const double *Lbase, int t, int c,
                                               uint32_t (&w)[G], double (&L)[G])
{
    const int last = c > 0 ? c - 1 : 0;
#pragma unroll
    for (int u = 0; u < G; u++) {
        const int tt = t + u * S < c ? t + u * S : last;
        w[u] = wbase[(int64_t)tt * 64];
        L[u] = Lbase[(int64_t)tt * 64];
    }
}

// The bonds of a lane's row whose neighbour slots lie below the staged tile's end T1, in row
// order from bond t, as groups of G: the words and lengths of the next group are loaded
// before the current group is evaluated, so their latency hides behind G bond evaluations.
// Returns the first bond not evaluated (past the tile, or c).  BOND(wd, L, t) evaluates one.
template <int G, int S, class BOND>
__device__ __forceinline__ int pdx_walk_tile(const uint16_t *wbase, const double *Lbase, int t, int c,
                                             int T1, BOND &&bond)
{
    if (t >= c) return t;
    uint32_t gw[G];
    double gL[G];
    pdx_load_group<G, S>(wbase, Lbase, t, c, gw, gL);
    while (true) {
        uint32_t nw[G];
        double nL[G];
        pdx_load_group<G, S>(wbase, Lbase, t + G * S, c, nw, nL);
        // the leading bonds of the group inside the row and the tile (slots increase along
        // the row, so every bond after the first one past T1 is past it too)
        int kk = 0;
#pragma unroll
        for (int u = 0; u < G; u++)
            if (kk == u && t + u * S < c && (int)(gw[u] & PDX_W16_SLOT) < T1) kk = u + 1;
#pragma unroll
        for (int u = 0; u < G; u++)
            if (u < kk) bond(gw[u], gL[u], t + u * S);
        t += kk * S;
        if (kk < G) return t;
#pragma unroll
        for (int u = 0; u < G; u++) {
            gw[u] = nw[u];
            gL[u] = nL[u];
        }
    }
}

// ----------------------------------------------------------------------------------------
// per-point tail: NaN flag (check_for_nans, nans.jl:10-18), acceleration and velocity
// (_update_acc_and_vel!), then the first half of the next step (update_vel_half!,
// apply_boundary_conditions!, update_disp_and_pos!).  Operands are loaded before the bond
// loop (tail_load) so that their latency overlaps it.
// ----------------------------------------------------------------------------------------
struct tail_ops {
    double u[3], vh[3], be[3];
    int32_t slot;
};

__device__ __forceinline__ void tail_load(const pdx_step_args &A, int64_t p, tail_ops &T)
{
    const int flags = A.flags;
    T.slot = -1;
#pragma unroll
    for (int d = 0; d < 3; d++) T.u[d] = T.vh[d] = T.be[d] = 0.0;
    if (flags & PDX_F_FORCE_ONLY) return;
#pragma unroll
    for (int d = 0; d < 3; d++) T.vh[d] = A.vh[3 * p + d];
    if (flags & PDX_F_HAS_NEXT) {
#pragma unroll
        for (int d = 0; d < 3; d++) T.u[d] = A.u[3 * p + d];
        T.slot = A.bcslot[p];
    }
    if (flags & PDX_F_BEXT) {
#pragma unroll
        for (int d = 0; d < 3; d++) T.be[d] = A.bext[3 * p + d];
    }
}

__device__ __forceinline__ void point_tail(const pdx_step_args &A, int64_t p, double4 ri,
                                           const double b[3], const tail_ops &T, double rho)
{
    const int flags = A.flags;
    if (isnan(b[0]) || isnan(b[1]) || isnan(b[2]))
        atomicMin(A.flags_out, (unsigned long long)A.step);
    if (flags & PDX_F_FORCE_ONLY) {
        const double s = (flags & PDX_F_SCALE_VOL) ? ri.w : 1.0;
#pragma unroll
        for (int d = 0; d < 3; d++) A.bint_out[3 * p + d] = (flags & PDX_F_SCALE_VOL) ? b[d] * s : b[d];
        return;
    }
    double x[3] = {ri.x, ri.y, ri.z};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int64_t k = 3 * p + d;
        const double ad = (b[d] + T.be[d]) / rho;
        const double vd = T.vh[d] + ad * A.dth;
        if (flags & PDX_F_STORE) {
            A.v_out[k] = vd;
            A.a_out[k] = ad;
            A.bint_out[k] = b[d];
        }
        if (flags & PDX_F_HAS_NEXT) {
            double vn = vd + ad * A.dth;
            if (T.slot >= 0) {
                const int32_t src = A.slot_src[3 * T.slot + d];
                if (src >= 0) {
                    const double val = A.bcv[src];
                    if (isfinite(val)) vn = val;
                }
            }
            const double du = vn * A.dt;
            A.u[k] = T.u[d] + du;
            x[d] = x[d] + du;
            A.vh[k] = vn;
        }
    }
    if (flags & PDX_F_HAS_NEXT) A.rec_out[p] = make_double4(x[0], x[1], x[2], ri.w);
}

// The same tail split over a point's four lanes (they all hold the point's sums): lane d < 3
// updates component d, lane 3 carries the record's volume, so that every load and store of
// the tail is one contiguous run per point.  Arithmetic as point_tail, component by
// component.
__device__ __forceinline__ void point_tail4(const pdx_step_args &A, int64_t p, double4 ri,
                                            const double b[3], int d, double rho)
{
    const int flags = A.flags;
    if (d == 0 && (isnan(b[0]) || isnan(b[1]) || isnan(b[2])))
        atomicMin(A.flags_out, (unsigned long long)A.step);
    const int64_t k = 3 * p + d;
    const double bd = d == 0 ? b[0] : (d == 1 ? b[1] : b[2]);
    if (flags & PDX_F_FORCE_ONLY) {
        if (d < 3) A.bint_out[k] = (flags & PDX_F_SCALE_VOL) ? bd * ri.w : bd;
        return;
    }
    double xo = d == 0 ? ri.x : (d == 1 ? ri.y : (d == 2 ? ri.z : ri.w));
    if (d < 3) {
        const double vh = A.vh[k];
        const double be = (flags & PDX_F_BEXT) ? A.bext[k] : 0.0;
        const double ad = (bd + be) / rho;
        const double vd = vh + ad * A.dth;
        if (flags & PDX_F_STORE) {
            A.v_out[k] = vd;
            A.a_out[k] = ad;
            A.bint_out[k] = bd;
        }
        if (flags & PDX_F_HAS_NEXT) {
            double vn = vd + ad * A.dth;
            const int32_t slot = A.bcslot[p];
            if (slot >= 0) {
                const int32_t src = A.slot_src[3 * slot + d];
                if (src >= 0) {
                    const double val = A.bcv[src];
                    if (isfinite(val)) vn = val;
                }
            }
            const double du = vn * A.dt;
            A.u[k] = A.u[k] + du;
            xo = xo + du;
            A.vh[k] = vn;
        }
    }
    if (flags & PDX_F_HAS_NEXT) reinterpret_cast<double *>(A.rec_out)[4 * p + d] = xo;
}

// sqrt(d2) correctly rounded for a normal, finite d2 > 0 (the Markstein sequence of the
// compiler's sqrt without its range scaling: bond lengths never come near the subnormal
// range), plus 1/sqrt(d2) to about 1 ulp as a by-product.
__device__ __forceinline__ double pdx_sqrt_rsq(double d2, double &rl)
{
    const double r = __builtin_amdgcn_rsq(d2);
    double g = d2 * r;
    double h = 0.5 * r;
    const double e = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, e, g);
    h = __builtin_fma(h, e, h);
    double d = __builtin_fma(-g, g, d2);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, d2);
    rl = h + h;
    return __builtin_fma(d, h, g);
}

// 1/x with one Newton step (relative error ~2^-50)
__device__ __forceinline__ double pdx_rcp1(double x)
{
    const double r = __builtin_amdgcn_rcp(x);
    return __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
}

// calc_failure! of one bond (bond_based.jl:125-144, ordinary_state_based.jl:107-126,
// bond_system.jl:309-327): breaks the bond (writes its word back) if eps > eps_c and the
// bond may fail.  eps = e * (1/L) is within ~1e-13 of the correctly rounded (l - L) / L;
// inside [epsc_lo, epsc_hi] the correctly rounded quotient decides, so a bond breaks
// exactly when the reference's breaks.
__device__ __forceinline__ bool bond_fail(const pdx_step_args &A, uint32_t *wp, uint32_t wd,
                                          double e, double Lb, double eps)
{
    bool act = (wd & PDX_BW_ACT) != 0;
    if (act && (wd & PDX_BW_FP) && eps >= A.epsc_lo) {
        const bool brk = (eps > A.epsc_hi) ? true : ((e / Lb) > A.epsc);
        if (brk) {
            act = false;
            *wp = wd & ~PDX_BW_ACT;
        }
    }
    return act;
}

// ----------------------------------------------------------------------------------------
// BB: calc_failure! (bond_based.jl:125-144) + calc_damage! (bond_system.jl:338-343) +
// force_density_point! (bond_based.jl:146-161) + the fused tail.  One thread per point, U
// bonds per iteration with all their loads and gathers issued before the arithmetic.
//
// FP64: l = sqrt((dx*dx + dy*dy) + dz*dz) is correctly rounded.  The force term is
// ((w bc) eps V_j) * (1/l), accumulated with dx by fused multiply-adds (the reference
// divides each component by l and adds).  An inactive bond has w = 0, as the reference's
// Bool * factor, so NaN positions still show.
// MIXED: the stretch comes from FP64 differences du = (x_j - x_i) - (X_j - X_i), then
// eps = (2 dX.du + du.du) / (L (l + L)) in FP32 and the force in FP32, summed in FP64.
// ----------------------------------------------------------------------------------------
template <bool MIXED, int U>
__global__ __launch_bounds__(PDX_BLOCK) void k_bb_step(pdx_step_args A)
{
    const int64_t p = pdx_thread_point();
    if (p >= A.n) return;
    const int c = A.cnt[p];
    const int64_t b0 = pdx_bond0(A.slice_ptr, p);
    uint32_t *const wbase = A.bw + b0;
    const double4 ri = A.rec_in[p];
    double4 Xi = make_double4(0, 0, 0, 0);
    if (MIXED) Xi = A.Xr[p];
    double bx = 0.0, by = 0.0, bz = 0.0;
    int na = 0;
    const double bc = A.bc;
    const float bcf = (float)A.bc, epscf = (float)A.epsc;
    for (int t = 0; t < c; t += U) {
        uint32_t wd[U];
        double Ld[U];
        float Lfl[U];
        double4 rj[U], Xj[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            wd[u] = 0;
            Ld[u] = 1.0;
            Lfl[u] = 1.0f;
            if (t + u < c) {
                const int64_t o = (int64_t)(t + u) * 64;
                wd[u] = wbase[o];
                if (MIXED) Lfl[u] = A.Lf[b0 + o];
                else Ld[u] = A.L[b0 + o];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            rj[u] = make_double4(0, 0, 0, 0);
            Xj[u] = rj[u];
            if (t + u < c) {
                const uint32_t j = wd[u] & PDX_BW_NBR;
                rj[u] = A.rec_in[j];
                if (MIXED) Xj[u] = A.Xr[j];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (t + u >= c) continue;
            const double dx = rj[u].x - ri.x, dy = rj[u].y - ri.y, dz = rj[u].z - ri.z;
            uint32_t *const wp = wbase + (int64_t)(t + u) * 64;
            if (!MIXED) {
                const double d2 = (dx * dx + dy * dy) + dz * dz;
                double rl;
                const double l = pdx_sqrt_rsq(d2, rl);
                const double e = l - Ld[u];
                const double eps = e * pdx_rcp1(Ld[u]);
                const bool act = bond_fail(A, wp, wd[u], e, Ld[u], eps);
                const double wbc = act ? bc : 0.0;  // (w * bc) with w = bond_active
                const double s = ((wbc * eps) * rj[u].w) * rl;
                bx = __builtin_fma(s, dx, bx);
                by = __builtin_fma(s, dy, by);
                bz = __builtin_fma(s, dz, bz);
                na += act ? 1 : 0;
            } else {
                const double DXd = Xj[u].x - Xi.x, DYd = Xj[u].y - Xi.y, DZd = Xj[u].z - Xi.z;
                const float ux = (float)(dx - DXd), uy = (float)(dy - DYd), uz = (float)(dz - DZd);
                const float DX = (float)DXd, DY = (float)DYd, DZ = (float)DZd;
                const float num = 2.0f * ((DX * ux + DY * uy) + DZ * uz) +
                                  ((ux * ux + uy * uy) + uz * uz);
                const float Lb = Lfl[u];
                const float l = __fsqrt_rn(Lb * Lb + num);
                const float eps = __fdividef(num, Lb * (l + Lb));
                bool act = (wd[u] & PDX_BW_ACT) != 0;
                if (act && (wd[u] & PDX_BW_FP) && eps > epscf) {
                    act = false;
                    *wp = wd[u] & ~PDX_BW_ACT;
                }
                const float wbc = act ? bcf : 0.0f;
                const float sf = __fdividef((wbc * eps) * (float)rj[u].w, l);
                bx += (double)(sf * (float)dx);
                by += (double)(sf * (float)dy);
                bz += (double)(sf * (float)dz);
                na += act ? 1 : 0;
            }
        }
    }
    if (A.flags & (PDX_F_STORE | PDX_F_FORCE_ONLY)) {
        A.nact_out[p] = na;
        A.dmg_out[p] = 1.0 - (double)na / (double)c;
    }
    tail_ops T;
    tail_load(A, p, T);
    const double b[3] = {bx, by, bz};
    point_tail(A, p, ri, b, T, A.rho);
}

// ----------------------------------------------------------------------------------------
// BB, LDS-staged (the production kernel).  A workgroup owns PDX_BLOCK consecutive Morton
// points.  It first copies the records of its halo (every point its bonds reach, ~2000
// for a 3.015-spacing horizon) from global memory into LDS with one coalesced pass, then
// each thread walks its point's bonds: 16-bit bond words (slot into the halo, active,
// fail_permit) and reference lengths stream from HBM, neighbour records come from LDS.
// Arithmetic as k_bb_step (FP64); MIXED takes the stretch from the FP64 difference
// d2 - L^2 (fused, one rounding) and does the rest in FP32 with FP64 sums.
// ----------------------------------------------------------------------------------------
__device__ __forceinline__ bool bond_fail16(uint16_t *wp, uint32_t wd, double e, double Lb, double eps,
                                            double epsc, double epsc_lo, double epsc_hi)
{
    bool act = (wd & PDX_W16_ACT) != 0;
    if (act && (wd & PDX_W16_FP) && eps >= epsc_lo) {
        const bool brk = (eps > epsc_hi) ? true : ((e / Lb) > epsc);
        if (brk) {
            act = false;
            *wp = (uint16_t)(wd & ~PDX_W16_ACT);
        }
    }
    return act;
}

// One bond of the FP64 BB step (bond_based.jl:125-161): failure against the point's eps_c,
// then the force term (w bc) eps V_j dx / l accumulated into b; bcb = the bond's bc,
// sc = its surface-correction factor (1 without).
// (rL = pdx_rcp1(Lb), passed in where one bond is evaluated at two states)
__device__ __forceinline__ void bb_bond64_r(double4 ri, double4 rj, uint16_t *wp, uint32_t wd, double Lb,
                                            double rL, double sc, double bcb, double epsc, double epsc_lo,
                                            double epsc_hi, double &bx, double &by, double &bz, int &na)
{
    const double dx = rj.x - ri.x, dy = rj.y - ri.y, dz = rj.z - ri.z;
    const double d2 = (dx * dx + dy * dy) + dz * dz;
    double rl;
    const double l = pdx_sqrt_rsq(d2, rl);
    const double e = l - Lb;
    const double eps = e * rL;
    const bool act = bond_fail16(wp, wd, e, Lb, eps, epsc, epsc_lo, epsc_hi);
    const double wbc = act ? sc * bcb : 0.0;  // (w * bc), w = active * sc
    const double s = ((wbc * eps) * rj.w) * rl;
    bx = __builtin_fma(s, dx, bx);
    by = __builtin_fma(s, dy, by);
    bz = __builtin_fma(s, dz, bz);
    na += act ? 1 : 0;
}

__device__ __forceinline__ void bb_bond64(double4 ri, double4 rj, uint16_t *wp, uint32_t wd, double Lb,
                                          double sc, double bcb, double epsc, double epsc_lo,
                                          double epsc_hi, double &bx, double &by, double &bz, int &na)
{
    bb_bond64_r(ri, rj, wp, wd, Lb, pdx_rcp1(Lb), sc, bcb, epsc, epsc_lo, epsc_hi, bx, by, bz, na);
}

// The same bond with one parameter set, whose neighbour record carries V_j bc in its fourth
// word (the neighbour-table staging forms it once per entry): one multiply fewer per bond.
__device__ __forceinline__ void bb_bond64_vbc(double4 ri, double4 rj, uint16_t *wp, uint32_t wd, double Lb,
                                              double rL, double sc, double epsc, double epsc_lo,
                                              double epsc_hi, double &bx, double &by, double &bz, int &na)
{
    const double dx = rj.x - ri.x, dy = rj.y - ri.y, dz = rj.z - ri.z;
    const double d2 = (dx * dx + dy * dy) + dz * dz;
    double rl;
    const double l = pdx_sqrt_rsq(d2, rl);
    const double e = l - Lb;
    const double eps = e * rL;
    const bool act = bond_fail16(wp, wd, e, Lb, eps, epsc, epsc_lo, epsc_hi);
    const double s = ((act ? sc * eps : 0.0) * rj.w) * rl;  // w eps (V_j bc) / l, w = active * sc
    bx = __builtin_fma(s, dx, bx);
    by = __builtin_fma(s, dy, by);
    bz = __builtin_fma(s, dz, bz);
    na += act ? 1 : 0;
}

// a neighbour record (x, V) with V_j bc in the fourth word, for bb_bond64_vbc
__device__ __forceinline__ double4 pdx_rec_vbc(double4 r, double bc) { return make_double4(r.x, r.y, r.z, r.w * bc); }

// F = feature mask.  PDX_FEAT_MULTI: several parameter sets (force_density_point! with a
// ParameterHandler, bond_based.jl:163-180): eps_c of the row's point, (bc_i + bc_j) / 2 per
// bond, rho_i.  PDX_FEAT_SC: EnergySurfaceCorrection, w = bond_active * scfactor[bond]
// (bond_based.jl:156, bond_system_corrections.jl:71-74).  PDX_FEAT_CAP: the table may be
// longer than the LDS budget; it is staged in tiles of HL entries.  Rows and table are both
// sorted by neighbour, so each lane meets the tiles in order: it walks its bonds while their
// slots lie in the staged tile and resumes there after the next tile is staged.
template <int U, int S, int F>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_bb_lds(pdx_step_args A)
{
    constexpr bool MULTI = (F & PDX_FEAT_MULTI) != 0;
    constexpr bool SC = (F & PDX_FEAT_SC) != 0;
    constexpr bool CAP = (F & PDX_FEAT_CAP) != 0;
    extern __shared__ double4 hrec[];
    const int64_t k = pdx_xcd_block();
    const int64_t p0 = k * PDX_BLOCK;
    if (p0 >= A.n) return;  // whole workgroup
    const int64_t h0 = A.halo_ptr[k];
    const int H = (int)(A.halo_ptr[k + 1] - h0);
    const int HL = CAP ? pdx_lds_entries(A.halo_max, A.lds_budget, 32 + (MULTI ? 1 : 0)) : A.halo_max;
    const int32_t *const hidx = A.halo + h0;
    uint8_t *const hpid = reinterpret_cast<uint8_t *>(hrec + HL);
    // S adjacent lanes share a point; lane s takes its bonds s, s + S, s + 2S, ...
    const int sub = threadIdx.x % S;
    const int64_t p = p0 + threadIdx.x / S;
    double bx = 0.0, by = 0.0, bz = 0.0;
    int na = 0;
    int c = 0;
    double4 ri = make_double4(0.0, 0.0, 0.0, 0.0);
    double rho = A.rho;
    if constexpr (!CAP) {
        // the whole table fits: stage it once
        for (int h = threadIdx.x; h < H; h += PDX_BLOCK * S) {
            const int32_t j = hidx[h];
            hrec[h] = MULTI ? A.rec_in[j] : pdx_rec_vbc(A.rec_in[j], A.bc);
            if (MULTI) hpid[h] = A.pid[j];
        }
        __syncthreads();
        if (p >= A.n) return;  // all S lanes of the point together
        c = A.cnt[p];
        const int64_t b0 = pdx_bond0(A.slice_ptr, p);
        uint16_t *const wbase = A.bw16 + b0;
        const double *const Lbase = A.L + b0;
        ri = A.rec_in[p];
        double bc = A.bc, epsc = A.epsc, epsc_lo = A.epsc_lo, epsc_hi = A.epsc_hi;
        if (MULTI) {
            const pdx_pset Pi = A.psets[A.pid[p]];
            bc = Pi.bc;
            epsc = Pi.epsc;
            epsc_lo = Pi.epsc_lo;
            epsc_hi = Pi.epsc_hi;
            rho = Pi.rho;
        }
        for (int t = sub; t < c; t += U * S) {
            uint32_t wd[U];
            double Ld[U], Sd[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                wd[u] = 0;
                Ld[u] = 1.0;
                Sd[u] = 1.0;
                if (t + u * S < c) {
                    const int64_t o = (int64_t)(t + u * S) * 64;
                    wd[u] = wbase[o];
                    Ld[u] = Lbase[o];
                    if (SC) Sd[u] = A.scf[b0 + o];
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (t + u * S >= c) continue;
                const int slot = (int)(wd[u] & PDX_W16_SLOT);
                if (MULTI)
                    bb_bond64(ri, hrec[slot], wbase + (int64_t)(t + u * S) * 64, wd[u], Ld[u], Sd[u],
                              (bc + A.psets[hpid[slot]].bc) / 2.0, epsc, epsc_lo, epsc_hi, bx, by, bz, na);
                else
                    bb_bond64_vbc(ri, hrec[slot], wbase + (int64_t)(t + u * S) * 64, wd[u], Ld[u],
                                  pdx_rcp1(Ld[u]), Sd[u], epsc, epsc_lo, epsc_hi, bx, by, bz, na);
            }
        }
    } else {
        // tiles of HL entries; all threads of the block take part in the staging barriers
        const bool valid = p < A.n;
        c = valid ? A.cnt[p] : 0;
        const int64_t b0 = valid ? pdx_bond0(A.slice_ptr, p) : 0;
        uint16_t *const wbase = A.bw16 + b0;
        const double *const Lbase = A.L + b0;
        if (valid) ri = A.rec_in[p];
        double bc = A.bc, epsc = A.epsc, epsc_lo = A.epsc_lo, epsc_hi = A.epsc_hi;
        if (MULTI && valid) {
            const pdx_pset Pi = A.psets[A.pid[p]];
            bc = Pi.bc;
            epsc = Pi.epsc;
            epsc_lo = Pi.epsc_lo;
            epsc_hi = Pi.epsc_hi;
            rho = Pi.rho;
        }
        int t = sub;
        for (int T0 = 0; T0 < H; T0 += HL) {
            const int T1 = H < T0 + HL ? H : T0 + HL;
            if (T0 > 0) __syncthreads();  // the previous tile is consumed
            for (int h = T0 + (int)threadIdx.x; h < T1; h += PDX_BLOCK * S) {
                const int32_t j = hidx[h];
                hrec[h - T0] = MULTI ? A.rec_in[j] : pdx_rec_vbc(A.rec_in[j], A.bc);
                if (MULTI) hpid[h - T0] = A.pid[j];
            }
            __syncthreads();
            for (; t < c; t += U * S) {
                uint32_t wd[U];
                double Ld[U], Sd[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    wd[u] = 0;
                    Ld[u] = 1.0;
                    Sd[u] = 1.0;
                    if (t + u * S < c) {
                        const int64_t o = (int64_t)(t + u * S) * 64;
                        wd[u] = wbase[o];
                        Ld[u] = Lbase[o];
                        if (SC) Sd[u] = A.scf[b0 + o];
                    }
                }
                // the leading bonds of the group whose neighbours lie in the staged tile
                int kk = 0;
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (kk == u && t + u * S < c && (int)(wd[u] & PDX_W16_SLOT) < T1) kk = u + 1;
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (u >= kk) continue;
                    const int ls = (int)(wd[u] & PDX_W16_SLOT) - T0;
                    if (MULTI)
                        bb_bond64(ri, hrec[ls], wbase + (int64_t)(t + u * S) * 64, wd[u], Ld[u], Sd[u],
                                  (bc + A.psets[hpid[ls]].bc) / 2.0, epsc, epsc_lo, epsc_hi, bx, by, bz, na);
                    else
                        bb_bond64_vbc(ri, hrec[ls], wbase + (int64_t)(t + u * S) * 64, wd[u], Ld[u],
                                      pdx_rcp1(Ld[u]), Sd[u], epsc, epsc_lo, epsc_hi, bx, by, bz, na);
                }
                if (kk < U) {
                    t += kk * S;  // resume at the first bond past the tile (or the row's end)
                    break;
                }
            }
        }
#pragma unroll
        for (int o = 1; o < S; o <<= 1) {
            bx += __shfl_xor(bx, o);
            by += __shfl_xor(by, o);
            bz += __shfl_xor(bz, o);
            na += __shfl_xor(na, o);
        }
        if (!valid) return;
    }
    if constexpr (!CAP) {
#pragma unroll
        for (int o = 1; o < S; o <<= 1) {
            bx += __shfl_xor(bx, o);
            by += __shfl_xor(by, o);
            bz += __shfl_xor(bz, o);
            na += __shfl_xor(na, o);
        }
    }
    if (sub == 0 && (A.flags & (PDX_F_STORE | PDX_F_FORCE_ONLY))) {
        A.nact_out[p] = na;
        A.dmg_out[p] = 1.0 - (double)na / (double)c;
    }
    static_assert(S == 4, "the tail takes four lanes per point");
    const double b[3] = {bx, by, bz};
    point_tail4(A, p, ri, b, sub, rho);
}

// ----------------------------------------------------------------------------------------
// J v of the bond-based material by central differences (jacobian_vector_product!,
// newton_krylov.jl:504-532) in one kernel: the block stages its neighbour table at the
// positions X + (u + e v), walks its bonds, stages the table again at X + (u - e v) and
// walks them again, and combines the two force densities (times the volume) into
// (F+ - F-) / (2 e); constrained rows take v.  The positions are formed as the reference
// forms them (u +- e v, then X + u, two roundings each), the bond arithmetic is that of
// k_bb_lds, so the result equals the two force evaluations of the step kernel bit for bit.
// Bonds that break at the + state stay broken at the - state, as in two successive
// calc_force_density! calls; n_active and damage keep the - state's values.
// ----------------------------------------------------------------------------------------
__device__ __forceinline__ double4 jv_position(const double4 X, const double *u, const double *v,
                                               int64_t j, double e)
{
    const double u0 = u[3 * j] + e * v[3 * j];
    const double u1 = u[3 * j + 1] + e * v[3 * j + 1];
    const double u2 = u[3 * j + 2] + e * v[3 * j + 2];
    return make_double4(X.x + u0, X.y + u1, X.z + u2, X.w);
}

template <int U, int S>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_bb_jv(pdx_step_args A, const double *u, const double *v,
                                                          const double *un2, const double *parts, int np,
                                                          double sc, const uint8_t *cmask, double *jv)
{
    extern __shared__ double4 hrec[];
    __shared__ double shs[PDX_BLOCK * S / 64];
    const int64_t k = pdx_xcd_block();
    const int64_t p0 = k * PDX_BLOCK;
    if (p0 >= A.n) return;  // whole workgroup
    const double vn2 = pdx_block_sum_parts(parts, np, shs);
    const int64_t h0 = A.halo_ptr[k];
    const int H = (int)(A.halo_ptr[k + 1] - h0);
    const int32_t *const hidx = A.halo + h0;
    const int sub = threadIdx.x % S;
    const int64_t p = p0 + threadIdx.x / S;
    const bool valid = p < A.n;
    const double e = pdx_nk_perturbation_v(*un2, vn2, sc);
    const int c = valid ? A.cnt[p] : 0;
    const int64_t b0 = valid ? pdx_bond0(A.slice_ptr, p) : 0;
    uint16_t *const wbase = A.bw16 + b0;
    const double *const Lbase = A.L + b0;
    // lane d < 3 keeps component d of F+ (the first pass); after the second, of F-
    double fpd = 0.0, fmd = 0.0;
    bool bad = false;
    int na = 0;
#pragma unroll 1
    for (int pass = 0; pass < 2; pass++) {
        const double es = pass == 0 ? e : -e;
        if (pass == 1) __syncthreads();  // the + table is consumed
        for (int h = threadIdx.x; h < H; h += PDX_BLOCK * S) {
            const int32_t j = hidx[h];
            hrec[h] = pdx_rec_vbc(jv_position(A.Xr[j], u, v, j, es), A.bc);
        }
        __syncthreads();
        if (!valid) continue;  // all S lanes of the point together
        const double4 ri = jv_position(A.Xr[p], u, v, p, es);
        double bx = 0.0, by = 0.0, bz = 0.0;
        int nap = 0;
        for (int t = sub; t < c; t += U * S) {
            uint32_t wd[U];
            double Ld[U];
#pragma unroll
            for (int q = 0; q < U; q++) {
                wd[q] = 0;
                Ld[q] = 1.0;
                if (t + q * S < c) {
                    const int64_t o = (int64_t)(t + q * S) * 64;
                    wd[q] = wbase[o];
                    Ld[q] = Lbase[o];
                }
            }
#pragma unroll
            for (int q = 0; q < U; q++) {
                if (t + q * S >= c) continue;
                const int slot = (int)(wd[q] & PDX_W16_SLOT);
                bb_bond64_vbc(ri, hrec[slot], wbase + (int64_t)(t + q * S) * 64, wd[q], Ld[q],
                              pdx_rcp1(Ld[q]), 1.0, A.epsc, A.epsc_lo, A.epsc_hi, bx, by, bz, nap);
            }
        }
#pragma unroll
        for (int o = 1; o < S; o <<= 1) {
            bx += __shfl_xor(bx, o);
            by += __shfl_xor(by, o);
            bz += __shfl_xor(bz, o);
            nap += __shfl_xor(nap, o);
        }
        bad = bad || isnan(bx) || isnan(by) || isnan(bz);
        const double fd = sub == 0 ? bx : (sub == 1 ? by : bz);
        if (pass == 0) fpd = fd;
        else fmd = fd;
        na = nap;
    }
    if (!valid) return;
    if (sub == 0) {
        A.nact_out[p] = na;
        A.dmg_out[p] = 1.0 - (double)na / (double)c;
        if (bad) atomicMin(A.flags_out, (unsigned long long)A.step);
    }
    if (sub < 3) {
        const int64_t kk = 3 * p + sub;
        const double w = A.Xr[p].w;
        jv[kk] = cmask[kk] ? v[kk] : (fpd * w - fmd * w) / (2.0 * e);
    }
}

// The same J v with both tables staged at once (twice the LDS of k_bb_jv, one block per CU
// for a typical table): each bond's word and length are read once and the bond is
// evaluated at both states.  A bond that breaks at the + state is evaluated broken at the
// - state, as above.
template <int U, int S>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_bb_jv1(pdx_step_args A, const double *u, const double *v,
                                                           const double *un2, const double *parts, int np,
                                                           double sc, const uint8_t *cmask, double *jv)
{
    extern __shared__ double4 hrec[];  // [0, H): + positions, [H, 2H): - positions
    __shared__ double shs[PDX_BLOCK * S / 64];
    const int64_t k = pdx_xcd_block();
    const int64_t p0 = k * PDX_BLOCK;
    if (p0 >= A.n) return;  // whole workgroup
    const double vn2 = pdx_block_sum_parts(parts, np, shs);
    const int64_t h0 = A.halo_ptr[k];
    const int H = (int)(A.halo_ptr[k + 1] - h0);
    const int32_t *const hidx = A.halo + h0;
    const int sub = threadIdx.x % S;
    const int64_t p = p0 + threadIdx.x / S;
    const bool valid = p < A.n;  // the same for the S lanes of a point
    const double e = pdx_nk_perturbation_v(*un2, vn2, sc);
    for (int h = threadIdx.x; h < H; h += PDX_BLOCK * S) {
        const int32_t j = hidx[h];
        const double4 X = A.Xr[j];
        hrec[h] = pdx_rec_vbc(jv_position(X, u, v, j, e), A.bc);
        hrec[H + h] = pdx_rec_vbc(jv_position(X, u, v, j, -e), A.bc);
    }
    __syncthreads();
    if (valid) {
        const int c = A.cnt[p];
        const int64_t b0 = pdx_bond0(A.slice_ptr, p);
        uint16_t *const wbase = A.bw16 + b0;
        const double *const Lbase = A.L + b0;
        const double4 Xi = A.Xr[p];
        const double4 rp = jv_position(Xi, u, v, p, e), rm = jv_position(Xi, u, v, p, -e);
        double px = 0.0, py = 0.0, pz = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
        int nap = 0, nam = 0;
        for (int t = sub; t < c; t += U * S) {
            uint32_t wd[U];
            double Ld[U];
#pragma unroll
            for (int q = 0; q < U; q++) {
                wd[q] = 0;
                Ld[q] = 1.0;
                if (t + q * S < c) {
                    const int64_t o = (int64_t)(t + q * S) * 64;
                    wd[q] = wbase[o];
                    Ld[q] = Lbase[o];
                }
            }
#pragma unroll
            for (int q = 0; q < U; q++) {
                if (t + q * S >= c) continue;
                const int slot = (int)(wd[q] & PDX_W16_SLOT);
                uint16_t *const wp = wbase + (int64_t)(t + q * S) * 64;
                const int before = nap;
                const double rL = pdx_rcp1(Ld[q]);
                bb_bond64_vbc(rp, hrec[slot], wp, wd[q], Ld[q], rL, 1.0, A.epsc, A.epsc_lo, A.epsc_hi,
                              px, py, pz, nap);
                const uint32_t wm = nap != before ? wd[q] : (wd[q] & ~(uint32_t)PDX_W16_ACT);
                bb_bond64_vbc(rm, hrec[H + slot], wp, wm, Ld[q], rL, 1.0, A.epsc, A.epsc_lo, A.epsc_hi,
                              mx, my, mz, nam);
            }
        }
#pragma unroll
        for (int o = 1; o < S; o <<= 1) {
            px += __shfl_xor(px, o);
            py += __shfl_xor(py, o);
            pz += __shfl_xor(pz, o);
            mx += __shfl_xor(mx, o);
            my += __shfl_xor(my, o);
            mz += __shfl_xor(mz, o);
            nam += __shfl_xor(nam, o);
        }
        if (sub == 0) {
            A.nact_out[p] = nam;
            A.dmg_out[p] = 1.0 - (double)nam / (double)c;
            if (isnan(px) || isnan(py) || isnan(pz) || isnan(mx) || isnan(my) || isnan(mz))
                atomicMin(A.flags_out, (unsigned long long)A.step);
        }
        if (sub < 3) {
            const int64_t kk = 3 * p + sub;
            const double a = sub == 0 ? px : (sub == 1 ? py : pz);
            const double b = sub == 0 ? mx : (sub == 1 ? my : mz);
            jv[kk] = cmask[kk] ? v[kk] : (a * Xi.w - b * Xi.w) / (2.0 * e);
        }
    }
}

// ----------------------------------------------------------------------------------------
// Mixed-precision BB step on the LDS-staged layout.  The stretch is formed without
// cancellation from the reference bond vector D = X_j - X_i and the displacement difference
// U = u_j - u_i (u = x - X of the positions the step receives):  l^2 - L^2 = 2 D.U + U.U,  eps = (l^2 - L^2) / (L (l + L)),  L = |D|,
// l = |D + U|.  Both are staged per block in FP32 relative to the block's first point
// (subtracted in FP64 first), so their rounding is relative to the block's extent, not to
// the absolute coordinates, and a rigid motion of the body costs no precision.  Every
// bond operation is FP32; the per-point sums are FP64 (each FP32 term is added with FP64
// rounding, the effect of a compensated FP32 sum).  Per bond only the 16-bit bond word is
// read: L comes from D.  The Velocity Verlet tail stays FP64.
// ----------------------------------------------------------------------------------------
template <int U, int S, int F>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_bb_lds_mx(pdx_step_args A)
{
    constexpr bool MULTI = (F & PDX_FEAT_MULTI) != 0;
    constexpr bool SC = (F & PDX_FEAT_SC) != 0;
    constexpr bool CAP = (F & PDX_FEAT_CAP) != 0;
    extern __shared__ float4 hmx[];  // [2 H]: (X - O, V), (u - u_O, 0)
    const int64_t k = pdx_xcd_block();
    const int64_t p0 = k * PDX_BLOCK;
    if (p0 >= A.n) return;  // whole workgroup
    const int64_t h0 = A.halo_ptr[k];
    const int H = (int)(A.halo_ptr[k + 1] - h0);
    // displacements are taken as x - X of the positions the step receives, so that the
    // perturbed positions of the Newton-Krylov products work like the Verlet state
    const double4 O = A.Xr[p0], xo = A.rec_in[p0];
    const double uo0 = xo.x - O.x, uo1 = xo.y - O.y, uo2 = xo.z - O.z;
    const int HL = CAP ? pdx_lds_entries(A.halo_max, A.lds_budget, 32 + (MULTI ? 1 : 0)) : A.halo_max;
    const int Hs = H < HL ? H : HL;
    const int32_t *const hidx = A.halo + h0;
    uint8_t *const hpid = reinterpret_cast<uint8_t *>(hmx + 2 * HL);
    for (int h = threadIdx.x; h < Hs; h += PDX_BLOCK * S) {
        const int64_t j = hidx[h];
        if (MULTI) hpid[h] = A.pid[j];
        const double4 X = A.Xr[j], x = A.rec_in[j];
        hmx[2 * h] = make_float4((float)(X.x - O.x), (float)(X.y - O.y), (float)(X.z - O.z), (float)X.w);
        hmx[2 * h + 1] = make_float4((float)((x.x - X.x) - uo0), (float)((x.y - X.y) - uo1),
                                     (float)((x.z - X.z) - uo2), 0.f);
    }
    __syncthreads();
    const int sub = threadIdx.x % S;
    const int64_t p = p0 + threadIdx.x / S;
    if (p >= A.n) return;  // all S lanes of the point together
    const int c = A.cnt[p];
    const int64_t b0 = pdx_bond0(A.slice_ptr, p);
    uint16_t *const wbase = A.bw16 + b0;
    const double4 ri = A.rec_in[p];
    const double4 Xi = A.Xr[p];
    const float xi = (float)(Xi.x - O.x), yi = (float)(Xi.y - O.y), zi = (float)(Xi.z - O.z);
    const float ui = (float)((ri.x - Xi.x) - uo0), vi = (float)((ri.y - Xi.y) - uo1),
                wi = (float)((ri.z - Xi.z) - uo2);
    double bx = 0.0, by = 0.0, bz = 0.0;
    int na = 0;
    float bcf = (float)A.bc, epscf = (float)A.epsc;
    double rho = A.rho, epsc = A.epsc;
    if (MULTI) {
        const pdx_pset Pi = A.psets[A.pid[p]];
        bcf = (float)Pi.bc;
        epscf = (float)Pi.epsc;
        epsc = Pi.epsc;
        rho = Pi.rho;
    }
    // FP32 stretches within 1e-4 of eps_c are decided on the FP64 stretch, so that a bond
    // breaks in the step in which the FP64 step would break it (the FP32 stretch is good to
    // ~1e-6 relative)
    const float win_lo = epscf * 0.9999f, win_hi = epscf * 1.0001f;
    for (int t = sub; t < c; t += U * S) {
        uint32_t wd[U];
        float sc[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            wd[u] = (t + u * S < c) ? wbase[(int64_t)(t + u * S) * 64] : 0u;
            sc[u] = (SC && t + u * S < c) ? A.scf32[b0 + (int64_t)(t + u * S) * 64] : 1.f;
        }
        // the group's bond forces are summed in FP32 (fused), then once into the FP64 sums
        float gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (t + u * S >= c) continue;
            const int slot = (int)(wd[u] & PDX_W16_SLOT);
            float4 a, b;
            int32_t jg = 0;
            if (!CAP || slot < HL) {
                a = hmx[2 * slot];
                b = hmx[2 * slot + 1];
            } else {  // past the staged entries: the same values from global memory
                jg = hidx[slot];
                const double4 X = A.Xr[jg], x = A.rec_in[jg];
                a = make_float4((float)(X.x - O.x), (float)(X.y - O.y), (float)(X.z - O.z), (float)X.w);
                b = make_float4((float)((x.x - X.x) - uo0), (float)((x.y - X.y) - uo1),
                                (float)((x.z - X.z) - uo2), 0.f);
            }
            const float Dx = a.x - xi, Dy = a.y - yi, Dz = a.z - zi;
            const float Ux = b.x - ui, Uy = b.y - vi, Uz = b.z - wi;
            const float L2 = __builtin_fmaf(Dz, Dz, __builtin_fmaf(Dy, Dy, Dx * Dx));
            const float DU = __builtin_fmaf(Dz, Uz, __builtin_fmaf(Dy, Uy, Dx * Ux));
            const float UU = __builtin_fmaf(Uz, Uz, __builtin_fmaf(Uy, Uy, Ux * Ux));
            const float num = __builtin_fmaf(2.f, DU, UU);  // l^2 - L^2
            const float Lb = __builtin_amdgcn_sqrtf(L2);
            const float l = __builtin_amdgcn_sqrtf(L2 + num);
            const float eps = num * __builtin_amdgcn_rcpf(Lb * (l + Lb));
            bool act = (wd[u] & PDX_W16_ACT) != 0;
            if (act && (wd[u] & PDX_W16_FP) && eps > win_lo) {
                bool brk = eps > win_hi;
                if (!brk) {
                    const double Lb = A.L[b0 + (int64_t)(t + u * S) * 64];
                    const double4 xj = A.rec_in[hidx[slot]];
                    const double dx = xj.x - ri.x, dy = xj.y - ri.y, dz = xj.z - ri.z;
                    double rl;
                    const double l64 = pdx_sqrt_rsq((dx * dx + dy * dy) + dz * dz, rl);
                    brk = ((l64 - Lb) / Lb) > epsc;
                }
                if (brk) {
                    act = false;
                    wbase[(int64_t)(t + u * S) * 64] = (uint16_t)(wd[u] & ~PDX_W16_ACT);
                }
            }
            const float bcb = MULTI ? (bcf + (float)A.psets[(!CAP || slot < HL) ? hpid[slot] : A.pid[jg]].bc) * 0.5f : bcf;
            const float sf = act ? (((SC ? sc[u] * bcb : bcb) * eps) * a.w) * __builtin_amdgcn_rcpf(l) : 0.f;
            gx = __builtin_fmaf(sf, Dx + Ux, gx);
            gy = __builtin_fmaf(sf, Dy + Uy, gy);
            gz = __builtin_fmaf(sf, Dz + Uz, gz);
            na += act ? 1 : 0;
        }
        bx += (double)gx;
        by += (double)gy;
        bz += (double)gz;
    }
#pragma unroll
    for (int o = 1; o < S; o <<= 1) {
        bx += __shfl_xor(bx, o);
        by += __shfl_xor(by, o);
        bz += __shfl_xor(bz, o);
        na += __shfl_xor(na, o);
    }
    if (sub == 0 && (A.flags & (PDX_F_STORE | PDX_F_FORCE_ONLY))) {
        A.nact_out[p] = na;
        A.dmg_out[p] = 1.0 - (double)na / (double)c;
    }
    static_assert(S == 4, "the tail takes four lanes per point");
    const double bb[3] = {bx, by, bz};
    point_tail4(A, p, ri, bb, sub, rho);
}

// ----------------------------------------------------------------------------------------
// Mixed-precision OSB (one parameter set): the two passes of k_osb_lds1/2 with the bond
// arithmetic of k_bb_lds_mx.  The stretch comes from the block-relative reference bond
// vector D and displacement difference U in FP32 (e = l - L = (2 D.U + U.U) / (l + L), no
// cancellation); the weighted volume, the dilatation sum and the force density are summed
// in FP64 from FP32 bond terms; c1, c2 and the Verlet tail are FP64.  Pass 1 stores
// (c1, c2); pass 2 forms both terms of a bond, (w/l)[(c2_i L + c1_i e) V_j + (c2 V)_j L +
// (c1 V)_j e], from 36-byte table entries (D, V), (U, (c1 V)_j), (c2 V)_j, so that pass 1
// carries no force sums.  Failure: FP32 stretches within 1e-4 of eps_c are decided on the
// FP64 stretch.
// ----------------------------------------------------------------------------------------
__device__ __forceinline__ void mx_stage_xu(const pdx_step_args &A, int64_t j, double4 O,
                                            double uo0, double uo1, double uo2, float4 &a, float4 &b)
{
    const double4 X = A.Xr[j], x = A.rec_in[j];
    a = make_float4((float)(X.x - O.x), (float)(X.y - O.y), (float)(X.z - O.z), (float)X.w);
    b = make_float4((float)((x.x - X.x) - uo0), (float)((x.y - X.y) - uo1), (float)((x.z - X.z) - uo2), 0.f);
}

template <int U, int S, int F>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_osb_lds1_mx(pdx_step_args A)
{
    static_assert((F & PDX_FEAT_MULTI) == 0, "mixed OSB takes one parameter set");
    constexpr bool SC = (F & PDX_FEAT_SC) != 0;
    constexpr bool CAP = (F & PDX_FEAT_CAP) != 0;
    extern __shared__ float4 hmx[];
    const int64_t k = pdx_xcd_block();
    const int64_t p0 = k * PDX_BLOCK;
    if (p0 >= A.n) return;  // whole workgroup
    const int64_t h0 = A.halo_ptr[k];
    const int H = (int)(A.halo_ptr[k + 1] - h0);
    const double4 O = A.Xr[p0], xo = A.rec_in[p0];
    const double uo0 = xo.x - O.x, uo1 = xo.y - O.y, uo2 = xo.z - O.z;
    const int HL = CAP ? pdx_lds_entries(A.halo_max, A.lds_budget, 32) : A.halo_max;
    const int Hs = H < HL ? H : HL;
    const int32_t *const hidx = A.halo + h0;
    for (int h = threadIdx.x; h < Hs; h += PDX_BLOCK * S) {
        float4 a, b;
        mx_stage_xu(A, hidx[h], O, uo0, uo1, uo2, a, b);
        hmx[2 * h] = a;
        hmx[2 * h + 1] = b;
    }
    __syncthreads();
    const int sub = threadIdx.x % S;
    const int64_t p = p0 + threadIdx.x / S;
    if (p >= A.n) return;  // all S lanes of the point together
    const int c = A.cnt[p];
    const int64_t b0 = pdx_bond0(A.slice_ptr, p);
    uint16_t *const wbase = A.bw16 + b0;
    const double4 ri = A.rec_in[p];
    const double4 Xi = A.Xr[p];
    const float xi = (float)(Xi.x - O.x), yi = (float)(Xi.y - O.y), zi = (float)(Xi.z - O.z);
    const float ui = (float)((ri.x - Xi.x) - uo0), vi = (float)((ri.y - Xi.y) - uo1),
                wi = (float)((ri.z - Xi.z) - uo2);
    double wvol = 0.0, Ssum = 0.0;
    int na = 0;
    const double epsc = A.epsc;
    const float epscf = (float)epsc, deltaf = (float)A.delta;
    const float win_lo = epscf * 0.9999f, win_hi = epscf * 1.0001f;
    for (int t = sub; t < c; t += U * S) {
        uint32_t wd[U];
        float sc[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            wd[u] = (t + u * S < c) ? wbase[(int64_t)(t + u * S) * 64] : 0u;
            sc[u] = (SC && t + u * S < c) ? A.scf32[b0 + (int64_t)(t + u * S) * 64] : 1.f;
        }
        // the group's sums in FP32 (fused), then once into the FP64 sums
        float gv = 0.f, gs = 0.f;
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (t + u * S >= c) continue;
            const int slot = (int)(wd[u] & PDX_W16_SLOT);
            float4 a, b;
            if (!CAP || slot < HL) {
                a = hmx[2 * slot];
                b = hmx[2 * slot + 1];
            } else {  // past the staged entries: the same values from global memory
                mx_stage_xu(A, hidx[slot], O, uo0, uo1, uo2, a, b);
            }
            const float Dx = a.x - xi, Dy = a.y - yi, Dz = a.z - zi;
            const float Ux = b.x - ui, Uy = b.y - vi, Uz = b.z - wi;
            const float L2 = __builtin_fmaf(Dz, Dz, __builtin_fmaf(Dy, Dy, Dx * Dx));
            const float DU = __builtin_fmaf(Dz, Uz, __builtin_fmaf(Dy, Uy, Dx * Ux));
            const float UU = __builtin_fmaf(Uz, Uz, __builtin_fmaf(Uy, Uy, Ux * Ux));
            const float num = __builtin_fmaf(2.f, DU, UU);  // l^2 - L^2
            const float Lb = __builtin_amdgcn_sqrtf(L2);
            const float l = __builtin_amdgcn_sqrtf(L2 + num);
            const float rL = __builtin_amdgcn_rcpf(Lb);
            const float e = num * __builtin_amdgcn_rcpf(l + Lb);
            const float eps = e * rL;
            bool act = (wd[u] & PDX_W16_ACT) != 0;
            if (act && (wd[u] & PDX_W16_FP) && eps > win_lo) {
                bool brk = eps > win_hi;
                if (!brk) {
                    const double L64 = A.L[b0 + (int64_t)(t + u * S) * 64];
                    const double4 xj = A.rec_in[hidx[slot]];
                    const double dx = xj.x - ri.x, dy = xj.y - ri.y, dz = xj.z - ri.z;
                    double rl;
                    const double l64 = pdx_sqrt_rsq((dx * dx + dy * dy) + dz * dz, rl);
                    brk = ((l64 - L64) / L64) > epsc;
                }
                if (brk) {
                    act = false;
                    wbase[(int64_t)(t + u * S) * 64] = (uint16_t)(wd[u] & ~PDX_W16_ACT);
                }
            }
            if (act) {
                float w = deltaf * rL;  // linear_kernel(delta, L)
                if (SC) w *= sc[u];
                const float wVL = (w * a.w) * Lb;
                gv = __builtin_fmaf(wVL, Lb, gv);
                gs = __builtin_fmaf(wVL, e, gs);
                na++;
            }
        }
        wvol += (double)gv;
        Ssum += (double)gs;
    }
#pragma unroll
    for (int o = 1; o < S; o <<= 1) {
        wvol += __shfl_xor(wvol, o);
        Ssum += __shfl_xor(Ssum, o);
        na += __shfl_xor(na, o);
    }
    if (sub != 0) return;
    double c1 = 0.0, c2 = 0.0;
    if (wvol != 0.0) {
        const double dil = (3.0 / wvol) * Ssum;
        c1 = (15.0 * A.G) / wvol;
        c2 = dil * ((3.0 * A.K) / wvol - c1 / 3.0);
    }
    reinterpret_cast<double2 *>(A.pstate)[p] = make_double2(c1, c2);
    if (A.flags & (PDX_F_STORE | PDX_F_FORCE_ONLY)) {
        A.nact_out[p] = na;
        A.dmg_out[p] = 1.0 - (double)na / (double)c;
    }
}

template <int U, int S, int F>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_osb_lds2_mx(pdx_step_args A)
{
    static_assert((F & PDX_FEAT_MULTI) == 0, "mixed OSB takes one parameter set");
    constexpr bool SC = (F & PDX_FEAT_SC) != 0;
    constexpr bool CAP = (F & PDX_FEAT_CAP) != 0;
    extern __shared__ float4 hmx[];
    const int64_t k = pdx_xcd_block();
    const int64_t p0 = k * PDX_BLOCK;
    if (p0 >= A.n) return;  // whole workgroup
    const int64_t h0 = A.halo_ptr[k];
    const int H = (int)(A.halo_ptr[k + 1] - h0);
    const double4 O = A.Xr[p0], xo = A.rec_in[p0];
    const double uo0 = xo.x - O.x, uo1 = xo.y - O.y, uo2 = xo.z - O.z;
    const int HL = CAP ? pdx_lds_entries(A.halo_max, A.lds_budget, 36) : A.halo_max;
    const int Hs = H < HL ? H : HL;
    const int32_t *const hidx = A.halo + h0;
    const double2 *const cst = reinterpret_cast<const double2 *>(A.pstate);
    float *const hc2 = reinterpret_cast<float *>(hmx + 2 * HL);
    for (int h = threadIdx.x; h < Hs; h += PDX_BLOCK * S) {
        const int64_t j = hidx[h];
        float4 a, b;
        mx_stage_xu(A, j, O, uo0, uo1, uo2, a, b);
        const double2 cj = cst[j];
        const double Vj = A.Xr[j].w;
        b.w = (float)(cj.x * Vj);
        hmx[2 * h] = a;
        hmx[2 * h + 1] = b;
        hc2[h] = (float)(cj.y * Vj);
    }
    __syncthreads();
    const int sub = threadIdx.x % S;
    const int64_t p = p0 + threadIdx.x / S;
    if (p >= A.n) return;  // all S lanes of the point together
    const int c = A.cnt[p];
    const int64_t b0 = pdx_bond0(A.slice_ptr, p);
    const uint16_t *const wbase = A.bw16 + b0;
    const double4 ri = A.rec_in[p];
    const double4 Xi = A.Xr[p];
    const float xi = (float)(Xi.x - O.x), yi = (float)(Xi.y - O.y), zi = (float)(Xi.z - O.z);
    const float ui = (float)((ri.x - Xi.x) - uo0), vi = (float)((ri.y - Xi.y) - uo1),
                wi = (float)((ri.z - Xi.z) - uo2);
    double bx = 0.0, by = 0.0, bz = 0.0;
    const double2 ci = cst[p];
    const float c1i = (float)ci.x, c2i = (float)ci.y;
    const float deltaf = (float)A.delta;
    for (int t = sub; t < c; t += U * S) {
        uint32_t wd[U];
        float sc[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            wd[u] = (t + u * S < c) ? wbase[(int64_t)(t + u * S) * 64] : 0u;
            sc[u] = (SC && t + u * S < c) ? A.scf32[b0 + (int64_t)(t + u * S) * 64] : 1.f;
        }
        // the group's bond forces are summed in FP32 (fused), then once into the FP64 sums
        float gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (t + u * S >= c) continue;
            if (!(wd[u] & PDX_W16_ACT)) continue;  // one set: both copies of the bond agree
            const int slot = (int)(wd[u] & PDX_W16_SLOT);
            float4 a, b;
            float c2Vj;
            if (!CAP || slot < HL) {
                a = hmx[2 * slot];
                b = hmx[2 * slot + 1];
                c2Vj = hc2[slot];
            } else {
                const int64_t j = hidx[slot];
                mx_stage_xu(A, j, O, uo0, uo1, uo2, a, b);
                const double2 cj = cst[j];
                b.w = (float)(cj.x * (double)A.Xr[j].w);
                c2Vj = (float)(cj.y * (double)A.Xr[j].w);
            }
            const float Dx = a.x - xi, Dy = a.y - yi, Dz = a.z - zi;
            const float Ux = b.x - ui, Uy = b.y - vi, Uz = b.z - wi;
            const float L2 = __builtin_fmaf(Dz, Dz, __builtin_fmaf(Dy, Dy, Dx * Dx));
            const float DU = __builtin_fmaf(Dz, Uz, __builtin_fmaf(Dy, Uy, Dx * Ux));
            const float UU = __builtin_fmaf(Uz, Uz, __builtin_fmaf(Uy, Uy, Ux * Ux));
            const float num = __builtin_fmaf(2.f, DU, UU);
            const float Lb = __builtin_amdgcn_sqrtf(L2);
            const float l = __builtin_amdgcn_sqrtf(L2 + num);
            const float e = num * __builtin_amdgcn_rcpf(l + Lb);
            float w = deltaf * __builtin_amdgcn_rcpf(Lb);
            if (SC) w *= sc[u];
            // own term (c2_i L + c1_i e) V_j plus reverse term (c2 V)_j L + (c1 V)_j e
            const float fi = __builtin_fmaf(c1i, e, c2i * Lb) * a.w;
            const float fj = __builtin_fmaf(b.w, e, c2Vj * Lb);
            const float sf = (w * __builtin_amdgcn_rcpf(l)) * (fi + fj);
            gx = __builtin_fmaf(sf, Dx + Ux, gx);
            gy = __builtin_fmaf(sf, Dy + Uy, gy);
            gz = __builtin_fmaf(sf, Dz + Uz, gz);
        }
        bx += (double)gx;
        by += (double)gy;
        bz += (double)gz;
    }
#pragma unroll
    for (int o = 1; o < S; o <<= 1) {
        bx += __shfl_xor(bx, o);
        by += __shfl_xor(by, o);
        bz += __shfl_xor(bz, o);
    }
    static_assert(S == 4, "the tail takes four lanes per point");
    const double bb[3] = {bx, by, bz};
    point_tail4(A, p, ri, bb, sub, A.rho);
}

// ----------------------------------------------------------------------------------------
// OSB pass 1: calc_failure! (ordinary_state_based.jl:107-126), calc_damage!,
// calc_weighted_volume (:176-189), calc_dilatation (:191-204) -> c1, c2 (:135-136).
// |dX|^2 is taken as L*L (the reference forms dot(dX, dX); they differ by <= 2 ulp).
// ----------------------------------------------------------------------------------------
__global__ __launch_bounds__(PDX_BLOCK) void k_osb_pass1(pdx_step_args A)
{
    const int64_t p = pdx_thread_point();
    if (p >= A.n) return;
    const int c = A.cnt[p];
    const int64_t b0 = pdx_bond0(A.slice_ptr, p);
    const double4 ri = A.rec_in[p];
    double wvol = 0.0, S = 0.0;
    int na = 0;
    for (int t = 0; t < c; t++) {
        const int64_t b = b0 + (int64_t)t * 64;
        const uint32_t wd = A.bw[b];
        const double4 rj = A.rec_in[wd & PDX_BW_NBR];
        const double dx = rj.x - ri.x, dy = rj.y - ri.y, dz = rj.z - ri.z;
        const double Lb = A.L[b];
        double rl;
        const double l = pdx_sqrt_rsq((dx * dx + dy * dy) + dz * dz, rl);
        const double e = l - Lb;
        const double rL = pdx_rcp1(Lb);
        const bool act = bond_fail(A, A.bw + b, wd, e, Lb, e * rL);
        if (act) {
            const double w = pdx_omega(A.kernel, A.delta, Lb, rL);
            wvol += (w * (Lb * Lb)) * rj.w;
            S += ((w * Lb) * e) * rj.w;
            na++;
        }
    }
    double c1 = 0.0, c2 = 0.0;
    if (wvol != 0.0) {
        const double dil = (3.0 / wvol) * S;
        c1 = (15.0 * A.G) / wvol;
        c2 = dil * ((3.0 * A.K) / wvol - c1 / 3.0);
    }
    reinterpret_cast<double2 *>(A.pstate)[p] = make_double2(c1, c2);
    if (A.flags & (PDX_F_STORE | PDX_F_FORCE_ONLY)) {
        A.nact_out[p] = na;
        A.dmg_out[p] = 1.0 - (double)na / (double)c;
    }
}

// OSB pass 2: force_density_point! (ordinary_state_based.jl:128-149) as a gather:
// b_i = sum_j (w/l) dx_ij [ (c2_i L + c1_i (l-L)) + (c2_j L + c1_j (l-L)) ] V_j  (the term
// point j scatters into i carries V_j, the volume of the point whose loop emits it)
__global__ __launch_bounds__(PDX_BLOCK) void k_osb_pass2(pdx_step_args A)
{
    const int64_t p = pdx_thread_point();
    if (p >= A.n) return;
    const int c = A.cnt[p];
    const int64_t b0 = pdx_bond0(A.slice_ptr, p);
    const double4 ri = A.rec_in[p];
    const double2 ci = reinterpret_cast<const double2 *>(A.pstate)[p];
    double bx = 0.0, by = 0.0, bz = 0.0;
    for (int t = 0; t < c; t++) {
        const int64_t b = b0 + (int64_t)t * 64;
        const uint32_t wd = A.bw[b];
        if (!(wd & PDX_BW_ACT)) continue;
        const uint32_t j = wd & PDX_BW_NBR;
        const double4 rj = A.rec_in[j];
        const double2 cj = reinterpret_cast<const double2 *>(A.pstate)[j];
        const double dx = rj.x - ri.x, dy = rj.y - ri.y, dz = rj.z - ri.z;
        const double Lb = A.L[b];
        double rl;
        const double l = pdx_sqrt_rsq((dx * dx + dy * dy) + dz * dz, rl);
        const double w = pdx_omega(A.kernel, A.delta, Lb, pdx_rcp1(Lb));
        const double e = l - Lb;
        const double fi = ci.y * Lb + ci.x * e;
        const double fj = cj.y * Lb + cj.x * e;
        const double s = (w * rl) * ((fi + fj) * rj.w);
        bx = __builtin_fma(s, dx, bx);
        by = __builtin_fma(s, dy, by);
        bz = __builtin_fma(s, dz, bz);
    }
    tail_ops T;
    tail_load(A, p, T);
    const double b[3] = {bx, by, bz};
    point_tail(A, p, ri, b, T, A.rho);
}

// OSB on the LDS-staged layout (four lanes per point, as the BB step).  Bond states live
// in the 16-bit words.  With one parameter set the bond force of i-j,
//   (w / l) [(c2_i L + c1_i e) + (c2_j L + c1_j e)] V_j dx,
// splits into an own part, c2_i A_i + c1_i B_i with A_i = sum (w/l) V_j L dx and
// B_i = sum (w/l) V_j e dx, which pass 1 sums while it forms c1_i, c2_i, and a reverse
// part, (w/l) [(c2 V)_j L + (c1 V)_j e] dx, which pass 2 gathers.  Pass 1 stages x|V
// (32 B per neighbour-table entry) and stores (c1 V, c2 V) and the own part per point;
// pass 2 stages x and (c1 V, c2 V) (40 B per entry), so a 2 M-point lattice's longest
// table fits the LDS budget.  With several sets c1, c2 depend on the bond: pass 1 stores
// (wvol, dilatation) and pass 2 stages x|V and those (48 B + the set index per entry).
template <int S, int F>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_osb_lds1(pdx_step_args A)
{
    constexpr bool MULTI = (F & PDX_FEAT_MULTI) != 0;
    constexpr bool SPLIT = !MULTI;
    constexpr bool SC = (F & PDX_FEAT_SC) != 0;
    constexpr bool CAP = (F & PDX_FEAT_CAP) != 0;
    extern __shared__ double4 hrec[];
    const int64_t k = pdx_xcd_block();
    const int64_t p0 = k * PDX_BLOCK;
    if (p0 >= A.n) return;  // whole workgroup
    const int64_t h0 = A.halo_ptr[k];
    const int H = (int)(A.halo_ptr[k + 1] - h0);
    // CAP: tables longer than the LDS budget are staged in tiles of HL entries (rows and
    // tables sorted by neighbour: each lane meets the tiles in order)
    const int HL = CAP ? pdx_lds_entries(A.halo_max, A.lds_budget, 32 + (MULTI ? 1 : 0)) : A.halo_max;
    const int32_t *const hidx = A.halo + h0;
    uint8_t *const hpid = reinterpret_cast<uint8_t *>(hrec + HL);
    auto stage = [&](int T0, int T1) {
        for (int h = T0 + (int)threadIdx.x; h < T1; h += PDX_BLOCK * S) {
            const int32_t j = hidx[h];
            // with the linear influence function the record carries delta V_j (omega V_j =
            // (delta V_j) / L)
            const double4 r = A.rec_in[j];
            hrec[h - T0] = (F & PDX_FEAT_KERN) ? r : make_double4(r.x, r.y, r.z, A.delta * r.w);
            if (MULTI) hpid[h - T0] = A.pid[j];
        }
    };
    const int sub = threadIdx.x % S;
    const int64_t p = p0 + threadIdx.x / S;
    const bool valid = p < A.n;
    if constexpr (!CAP) {
        stage(0, H);
        __syncthreads();
        if (!valid) return;  // all S lanes of the point together
    }
    const int c = valid ? A.cnt[p] : 0;
    const int64_t b0 = valid ? pdx_bond0(A.slice_ptr, p) : 0;
    uint16_t *const wbase = A.bw16 + b0;
    const double *const Lbase = A.L + b0;
    const double4 ri = valid ? A.rec_in[p] : make_double4(0.0, 0.0, 0.0, 0.0);
    double wvol = 0.0, Ssum = 0.0;
    double ax = 0.0, ay = 0.0, az = 0.0, bx = 0.0, by = 0.0, bz = 0.0;  // A_i, B_i (SPLIT)
    int na = 0;
    double epsc = A.epsc, epsc_lo = A.epsc_lo, epsc_hi = A.epsc_hi, G = A.G, K = A.K;
    if (MULTI && valid) {
        const pdx_pset Pi = A.psets[A.pid[p]];
        epsc = Pi.epsc;
        epsc_lo = Pi.epsc_lo;
        epsc_hi = Pi.epsc_hi;
    }
    // two bonds per lane in flight: their words and lengths are loaded before either is used
    constexpr int U = 2;
    int t0 = sub;
    // the lane's bonds whose neighbours lie in the staged entries [T0, T1)
    auto walk = [&](int T0, int T1) {
        for (; t0 < c; t0 += U * S) {
            uint32_t wdu[U];
            double Lu[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const bool ok = t0 + u * S < c;
                wdu[u] = ok ? wbase[(int64_t)(t0 + u * S) * 64] : 0u;
                Lu[u] = ok ? Lbase[(int64_t)(t0 + u * S) * 64] : 1.0;
            }
            int kk = U;
            if (CAP) {
                kk = 0;
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (kk == u && t0 + u * S < c && (int)(wdu[u] & PDX_W16_SLOT) < T1) kk = u + 1;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int t = t0 + u * S;
                if (u >= kk || t >= c) continue;
                const int64_t o = (int64_t)t * 64;
                const uint32_t wd = wdu[u];
                const double Lb = Lu[u];
                const int ls = (int)(wd & PDX_W16_SLOT) - T0;
                const double4 rj = hrec[ls];
                const double dx = rj.x - ri.x, dy = rj.y - ri.y, dz = rj.z - ri.z;
                double rl;
                const double l = pdx_sqrt_rsq((dx * dx + dy * dy) + dz * dz, rl);
                const double e = l - Lb;
                const double rL = pdx_rcp1(Lb);
                bool act;
                if (MULTI) {
                    // The neighbour fails its own copy of the bond against its eps_c; the stretch
                    // is the same number at both ends, so the owner tracks that copy's state
                    // (RACT) exactly, which the gather of pass 2 needs.
                    const double eps = e * rL;
                    const pdx_pset Pj = A.psets[hpid[ls]];
                    uint32_t nw = wd;
                    if ((wd & PDX_W16_ACT) && (wd & PDX_W16_FP) && eps >= epsc_lo)
                        if ((eps > epsc_hi) ? true : ((e / Lb) > epsc)) nw &= ~PDX_W16_ACT;
                    if ((wd & PDX_W16_RACT) && (wd & PDX_W16_FP) && eps >= Pj.epsc_lo)
                        if ((eps > Pj.epsc_hi) ? true : ((e / Lb) > Pj.epsc)) nw &= ~PDX_W16_RACT;
                    if (nw != wd) wbase[o] = (uint16_t)nw;
                    act = (nw & PDX_W16_ACT) != 0;
                } else {
                    act = bond_fail16(wbase + o, wd, e, Lb, e * rL, epsc, epsc_lo, epsc_hi);
                }
                if (act) {
                    // omega, or 1 / L with delta already in the record
                    double w = (F & PDX_FEAT_KERN) ? pdx_omega(A.kernel, A.delta, Lb, rL) : rL;
                    if (SC) w *= A.scf[b0 + o];  // omega * beta
                    // w V_j formed once: wvol += (w V_j L) L, Ssum += (w V_j L) e
                    const double wV = w * rj.w;
                    const double wVL = wV * Lb;
                    wvol = __builtin_fma(wVL, Lb, wvol);
                    Ssum = __builtin_fma(wVL, e, Ssum);
                    na++;
                    if (SPLIT) {
                        const double q = wV * rl;
                        const double qL = q * Lb, qe = q * e;
                        ax = __builtin_fma(qL, dx, ax);
                        ay = __builtin_fma(qL, dy, ay);
                        az = __builtin_fma(qL, dz, az);
                        bx = __builtin_fma(qe, dx, bx);
                        by = __builtin_fma(qe, dy, by);
                        bz = __builtin_fma(qe, dz, bz);
                    }
                }
            }
            if (CAP && kk < U) {
                t0 += kk * S;  // resume at the first bond past the tile (or the row's end)
                return;
            }
        }
    };
    if constexpr (!CAP) {
        walk(0, H);
    } else {
        for (int T0 = 0; T0 < H; T0 += HL) {
            const int T1 = H < T0 + HL ? H : T0 + HL;
            if (T0 > 0) __syncthreads();  // the previous tile is consumed
            stage(T0, T1);
            __syncthreads();
            walk(T0, T1);
        }
    }
#pragma unroll
    for (int o = 1; o < S; o <<= 1) {
        wvol += __shfl_xor(wvol, o);
        Ssum += __shfl_xor(Ssum, o);
        na += __shfl_xor(na, o);
        if (SPLIT) {
            ax += __shfl_xor(ax, o);
            ay += __shfl_xor(ay, o);
            az += __shfl_xor(az, o);
            bx += __shfl_xor(bx, o);
            by += __shfl_xor(by, o);
            bz += __shfl_xor(bz, o);
        }
    }
    if (!valid) return;
    if (SPLIT) {
        // every lane holds the sums; lane d < 3 stores component d of the own part, lane 3
        // the per-point pair and the counts
        double c1 = 0.0, c2 = 0.0;
        if (wvol != 0.0) {
            const double dil = (3.0 / wvol) * Ssum;
            c1 = (15.0 * G) / wvol;
            c2 = dil * ((3.0 * K) / wvol - c1 / 3.0);
        }
        if (sub < 3) {
            const double a = sub == 0 ? ax : (sub == 1 ? ay : az);
            const double bb = sub == 0 ? bx : (sub == 1 ? by : bz);
            A.pstate[2 * A.n + 3 * p + sub] = c2 * a + c1 * bb;
            return;
        }
        reinterpret_cast<double2 *>(A.pstate)[p] = make_double2(c1 * ri.w, c2 * ri.w);
        if (A.flags & (PDX_F_STORE | PDX_F_FORCE_ONLY)) {
            A.nact_out[p] = na;
            A.dmg_out[p] = 1.0 - (double)na / (double)c;
        }
        return;
    }
    if (sub != 0) return;
    double c1 = 0.0, c2 = 0.0;
    if (MULTI) {
        // several sets: c1, c2 depend on the bond (averaged G, K); keep wvol and the dilatation
        c1 = wvol;
        c2 = wvol != 0.0 ? (3.0 / wvol) * Ssum : 0.0;
    } else if (wvol != 0.0) {
        const double dil = (3.0 / wvol) * Ssum;
        c1 = (15.0 * G) / wvol;
        c2 = dil * ((3.0 * K) / wvol - c1 / 3.0);
    }
    reinterpret_cast<double2 *>(A.pstate)[p] = make_double2(c1, c2);
    if (A.flags & (PDX_F_STORE | PDX_F_FORCE_ONLY)) {
        A.nact_out[p] = na;
        A.dmg_out[p] = 1.0 - (double)na / (double)c;
    }
}

// One set: the own part comes from pass 1, the loop gathers the reverse parts.
// MULTI: several parameter sets (force_density_point! with a ParameterHandler,
// ordinary_state_based.jl:151-174): pass 1 left (wvol, dilatation) per point and the bond
// takes c1 = 15 (G_i + G_j) / (2 wvol), c2 = dil (3 (K_i + K_j) / (2 wvol) - c1 / 3) of each
// end; a point with wvol = 0 contributes nothing (its loop returns early)
template <int S, int F>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_osb_lds2(pdx_step_args A)
{
    constexpr bool MULTI = (F & PDX_FEAT_MULTI) != 0;
    constexpr bool SPLIT = !MULTI;
    constexpr bool SC = (F & PDX_FEAT_SC) != 0;
    constexpr bool CAP = (F & PDX_FEAT_CAP) != 0;
    extern __shared__ double4 hrec[];
    // CAP: tables longer than the LDS budget are staged in tiles of HL entries
    const int HL = CAP ? pdx_lds_entries(A.halo_max, A.lds_budget, SPLIT ? 40 : 48 + 1) : A.halo_max;
    // SPLIT: hrec holds (x, y, z, (c1 V)_j), hc2 (c2 V)_j; otherwise hrec x|V, hc (wvol, dil)
    double2 *const hc = reinterpret_cast<double2 *>(hrec + HL);
    double *const hc2 = reinterpret_cast<double *>(hrec + HL);
    const int64_t k = pdx_xcd_block();
    const int64_t p0 = k * PDX_BLOCK;
    if (p0 >= A.n) return;  // whole workgroup
    const int64_t h0 = A.halo_ptr[k];
    const int H = (int)(A.halo_ptr[k + 1] - h0);
    const double2 *const cst = reinterpret_cast<const double2 *>(A.pstate);
    const int32_t *const hidx = A.halo + h0;
    uint8_t *const hpid = reinterpret_cast<uint8_t *>(hc + HL);
    auto stage = [&](int T0, int T1) {
        for (int h = T0 + (int)threadIdx.x; h < T1; h += PDX_BLOCK * S) {
            const int64_t j = hidx[h];
            const double4 r = A.rec_in[j];
            const double2 cj = cst[j];
            if (SPLIT) {
                // with the linear influence function: delta (c1 V)_j, delta (c2 V)_j
                const double dl = (F & PDX_FEAT_KERN) ? 1.0 : A.delta;
                hrec[h - T0] = make_double4(r.x, r.y, r.z, dl * cj.x);
                hc2[h - T0] = dl * cj.y;
            } else {
                hrec[h - T0] = r;
                hc[h - T0] = cj;
                hpid[h - T0] = A.pid[j];
            }
        }
    };
    const int sub = threadIdx.x % S;
    const int64_t p = p0 + threadIdx.x / S;
    const bool valid = p < A.n;
    if constexpr (!CAP) {
        stage(0, H);
        __syncthreads();
        if (!valid) return;  // all S lanes of the point together
    }
    const int c = valid ? A.cnt[p] : 0;
    const int64_t b0 = valid ? pdx_bond0(A.slice_ptr, p) : 0;
    const uint16_t *const wbase = A.bw16 + b0;
    const double *const Lbase = A.L + b0;
    const double4 ri = valid ? A.rec_in[p] : make_double4(0.0, 0.0, 0.0, 0.0);
    const double2 ci = (MULTI && valid) ? cst[p] : make_double2(0.0, 0.0);
    double bx = 0.0, by = 0.0, bz = 0.0;
    if (SPLIT && sub == 0 && valid) {
        const double *own = A.pstate + 2 * A.n + 3 * p;
        bx = own[0];
        by = own[1];
        bz = own[2];
    }
    double Gi = A.G, Ki = A.K, rho = A.rho;
    if (MULTI && valid) {
        const pdx_pset Pi = A.psets[A.pid[p]];
        Gi = Pi.G;
        Ki = Pi.K;
        rho = Pi.rho;
    }
    constexpr int U = 2;  // two bonds per lane in flight
    int t0 = sub;
    // the lane's bonds whose neighbours lie in the staged entries [T0, T1)
    auto walk = [&](int T0, int T1) {
        for (; t0 < c; t0 += U * S) {
            uint32_t wdu[U];
            double Lu[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const bool ok = t0 + u * S < c;
                wdu[u] = ok ? wbase[(int64_t)(t0 + u * S) * 64] : 0u;
                Lu[u] = ok ? Lbase[(int64_t)(t0 + u * S) * 64] : 1.0;
            }
            int kk = U;
            if (CAP) {
                kk = 0;
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (kk == u && t0 + u * S < c && (int)(wdu[u] & PDX_W16_SLOT) < T1) kk = u + 1;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int t = t0 + u * S;
                if (u >= kk || t >= c) continue;
                const int64_t o = (int64_t)t * 64;
                const uint32_t wd = wdu[u];
                const double Lb = Lu[u];
                const bool act = (wd & PDX_W16_ACT) != 0;
                const bool ract = MULTI ? (wd & PDX_W16_RACT) != 0 : act;  // one set: both copies agree
                if (!act && !ract) continue;
                const int ls = (int)(wd & PDX_W16_SLOT) - T0;
                double4 rj;
                double2 cj;
                if (SPLIT) {
                    rj = hrec[ls];
                    cj = make_double2(rj.w, hc2[ls]);
                } else {
                    rj = hrec[ls];
                    cj = hc[ls];
                }
                const double dx = rj.x - ri.x, dy = rj.y - ri.y, dz = rj.z - ri.z;
                double rl;
                const double l = pdx_sqrt_rsq((dx * dx + dy * dy) + dz * dz, rl);
                const double rLb = pdx_rcp1(Lb);
                // omega, or 1 / L where delta is already in the staged constants (one set)
                double w = (F & PDX_FEAT_KERN) ? pdx_omega(A.kernel, A.delta, Lb, rLb) : (SPLIT ? rLb : A.delta * rLb);
                if (SC) w *= A.scf[b0 + o];  // beta is symmetric: both ends' copies of the bond carry it
                const double e = l - Lb;
                if (SPLIT) {
                    // the reverse part (c2 V)_j L + (c1 V)_j e
                    const double s = (w * rl) * __builtin_fma(cj.x, e, cj.y * Lb);
                    bx = __builtin_fma(s, dx, bx);
                    by = __builtin_fma(s, dy, by);
                    bz = __builtin_fma(s, dz, bz);
                    continue;
                }
                double fi, fj;
                if (MULTI) {
                    const pdx_pset Pj = A.psets[hpid[ls]];
                    const double Gs = Gi + Pj.G, Ks = Ki + Pj.K;
                    fi = 0.0;
                    fj = 0.0;
                    if (act && ci.x != 0.0) {
                        const double c1 = (15.0 * Gs) / (2.0 * ci.x);
                        const double c2 = ci.y * ((3.0 * Ks) / (2.0 * ci.x) - c1 / 3.0);
                        fi = c2 * Lb + c1 * e;
                    }
                    if (ract && cj.x != 0.0) {
                        const double c1 = (15.0 * Gs) / (2.0 * cj.x);
                        const double c2 = cj.y * ((3.0 * Ks) / (2.0 * cj.x) - c1 / 3.0);
                        fj = c2 * Lb + c1 * e;
                    }
                } else {
                    fi = ci.y * Lb + ci.x * e;
                    fj = cj.y * Lb + cj.x * e;
                }
                const double s = (w * rl) * ((fi + fj) * rj.w);
                bx = __builtin_fma(s, dx, bx);
                by = __builtin_fma(s, dy, by);
                bz = __builtin_fma(s, dz, bz);
            }
            if (CAP && kk < U) {
                t0 += kk * S;  // resume at the first bond past the tile (or the row's end)
                return;
            }
        }
    };
    if constexpr (!CAP) {
        walk(0, H);
    } else {
        for (int T0 = 0; T0 < H; T0 += HL) {
            const int T1 = H < T0 + HL ? H : T0 + HL;
            if (T0 > 0) __syncthreads();  // the previous tile is consumed
            stage(T0, T1);
            __syncthreads();
            walk(T0, T1);
        }
    }
#pragma unroll
    for (int o = 1; o < S; o <<= 1) {
        bx += __shfl_xor(bx, o);
        by += __shfl_xor(by, o);
        bz += __shfl_xor(bz, o);
    }
    if (!valid) return;
    static_assert(S == 4, "the tail takes four lanes per point");
    const double b[3] = {bx, by, bz};
    point_tail4(A, p, ri, b, sub, rho);
}

// ----------------------------------------------------------------------------------------
// CPD (CMaterial, ZEMSilling).  3x3 matrices column-major as Julia: M[r + 3c].
// ----------------------------------------------------------------------------------------
__device__ __forceinline__ void m3_mul(const double *A, const double *B, double *C)
{
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++)
            C[r + 3 * c] = (A[r] * B[3 * c] + A[r + 3] * B[1 + 3 * c]) + A[r + 6] * B[2 + 3 * c];
}

__device__ __forceinline__ void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ double dot3(const double *a, const double *b)
{
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

__device__ __forceinline__ double m3_det(const double *A)
{
    double y[3];
    cross3(A + 3, A + 6, y);
    return dot3(A, y);
}

// inverse as adjugate rows (x1 x x2, x2 x x0, x0 x x1) over the determinant
__device__ __forceinline__ void m3_inv(const double *A, double *R)
{
    double x0[3] = {A[0], A[1], A[2]}, x1[3] = {A[3], A[4], A[5]}, x2[3] = {A[6], A[7], A[8]};
    double y0[3], y1[3], y2[3];
    cross3(x1, x2, y0);
    const double d = dot3(x0, y0);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        x0[k] = x0[k] / d;
        y0[k] = y0[k] / d;
    }
    cross3(x2, x0, y1);
    cross3(x0, x1, y2);
    R[0] = y0[0]; R[1] = y1[0]; R[2] = y2[0];
    R[3] = y0[1]; R[4] = y1[1]; R[5] = y2[1];
    R[6] = y0[2]; R[7] = y1[2]; R[8] = y2[2];
}

// first_piola_kirchhoff: SaintVenantKirchhoff (constitutive_models.jl:35-41) and
// NeoHookePenalty (:231-243)
// first_piola_kirchhoff of the constitutive models (constitutive_models.jl): P = F S
__device__ void pk1(int model, const double *F, double lambda, double mu, double G, double K,
                    double nu, double *P)
{
    double C[9], S[9];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++)
            C[r + 3 * c] = (F[3 * r] * F[3 * c] + F[1 + 3 * r] * F[1 + 3 * c]) + F[2 + 3 * r] * F[2 + 3 * c];
    if (model == PDX_NHP) {
        const double J = m3_det(F);
        if (J < 2.220446049250313e-16 || isnan(J)) {
            for (int k = 0; k < 9; k++) P[k] = 0.0;
            return;
        }
        double Ci[9];
        m3_inv(C, Ci);
        const double trC = (C[0] + C[4]) + C[8];
        const double a = (1.0 / 3.0) * trC;
        const double Jp = pow(J, -2.0 / 3.0);
        const double iJ = 1.0 / J;
        const double vol = (K / 4.0) * (J * J - iJ * iJ);
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const double Iab = (k % 4 == 0) ? 1.0 : 0.0;
            S[k] = (G * (Iab - a * Ci[k])) * Jp + vol * Ci[k];
        }
    } else if (model == PDX_NEO_HOOKE) {
        // :169-175: S = mu (I - C^-1) + lambda log(J) C^-1
        const double J = m3_det(F);
        double Ci[9];
        m3_inv(C, Ci);
        const double lJ = lambda * log(J);
#pragma unroll
        for (int k = 0; k < 9; k++) S[k] = mu * (((k % 4 == 0) ? 1.0 : 0.0) - Ci[k]) + lJ * Ci[k];
    } else if (model == PDX_LINEAR_ELASTIC) {
        // :83-97: S = C_voigt E_voigt, E_voigt = (E11, E22, E33, 2 E23, 2 E31, 2 E12),
        // C_voigt with a = (1 - nu) lambda / nu on the diagonal block and mu for shear
        double E[9];
#pragma unroll
        for (int k = 0; k < 9; k++) E[k] = 0.5 * (C[k] - ((k % 4 == 0) ? 1.0 : 0.0));
        const double a = ((1.0 - nu) * lambda) / nu;
        const double s1 = (a * E[0] + lambda * E[4]) + lambda * E[8];
        const double s2 = (lambda * E[0] + a * E[4]) + lambda * E[8];
        const double s3 = (lambda * E[0] + lambda * E[4]) + a * E[8];
        const double s4 = mu * (2.0 * E[7]), s5 = mu * (2.0 * E[2]), s6 = mu * (2.0 * E[3]);
        S[0] = s1; S[1] = s6; S[2] = s5;
        S[3] = s6; S[4] = s2; S[5] = s4;
        S[6] = s5; S[7] = s4; S[8] = s3;
    } else {
        double E[9];
#pragma unroll
        for (int k = 0; k < 9; k++) E[k] = 0.5 * (C[k] - ((k % 4 == 0) ? 1.0 : 0.0));
        const double lt = lambda * ((E[0] + E[4]) + E[8]);
        const double tm = 2.0 * mu;
#pragma unroll
        for (int k = 0; k < 9; k++) S[k] = tm * E[k] + ((k % 4 == 0) ? lt : 0.0);
    }
    m3_mul(F, S, P);
}

// The per-point part of CPD pass 1 (FP64): F = Ft K^-1, too_much_damage!, P, and
// Q = P K^-1 - beta F stored for pass 2; defgrad, cauchy_stress, damage and the active-bond
// count on export steps.  Km = K's six distinct entries (00 10 20 11 21 22).
template <bool ZW = false>
__device__ __forceinline__ bool cpd_point_state(const pdx_step_args &A, int64_t p, const double Km[6],
                                               const double Ft[9], int na, int c, bool dead,
                                               double lambda, double mu, double G, double K,
                                               double nu, double beta)
{
    const double K9[9] = {Km[0], Km[1], Km[2], Km[1], Km[3], Km[4], Km[2], Km[4], Km[5]};
    double Kinv[9], F[9];
    m3_inv(K9, Kinv);
    m3_mul(Ft, Kinv, F);
    const double dmg = 1.0 - (double)na / (double)c;
    bool hasnan = false;
#pragma unroll
    for (int k = 0; k < 9; k++) hasnan |= isnan(F[k]);
    const bool kill = dead || dmg > A.maxdmg || hasnan;
    const bool store = (A.flags & (PDX_F_STORE | PDX_F_FORCE_ONLY)) != 0;
    double PK[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (!kill) {
        double P[9];
        pk1(A.model, F, lambda, mu, G, K, nu, P);
        m3_mul(P, Kinv, PK);
        if (store) {
            // cauchy_stress: stress.jl:112-116, sigma = (1/J) P F'
            const double sj = 1.0 / m3_det(F);
#pragma unroll
            for (int k = 0; k < 9; k++) {
                const int r = k % 3, cc = k / 3;
                A.cauchy[9 * p + k] =
                    ((sj * P[r]) * F[cc] + (sj * P[r + 3]) * F[cc + 3]) + (sj * P[r + 6]) * F[cc + 6];
            }
        }
    }
    if (ZW) {
        // ZEMWan (correspondence.jl:253-285): tzem = w C_1 z with C_1 = C : K^-1 of the
        // point's Hooke tensor (zem_stabilization.jl:32-42; get_hooke_matrix, Voigt diagonal
        // a = (1 - nu) lambda / nu, lambda off it, mu for shear), so Q = P K^-1 - C_1 F and
        // pass 2 also needs C_1
        const double a = ((1.0 - nu) * lambda) / nu;
        double C1[9], CF[9];
#pragma unroll
        for (int cc = 0; cc < 3; cc++)
#pragma unroll
            for (int r = 0; r < 3; r++) {
                if (r == cc) {
                    const double d0 = r == 0 ? a : lambda, d1 = r == 1 ? a : lambda, d2 = r == 2 ? a : lambda;
                    C1[r + 3 * cc] = (d0 * Kinv[0] + d1 * Kinv[4]) + d2 * Kinv[8];
                } else {
                    C1[r + 3 * cc] = mu * Kinv[r + 3 * cc] + mu * Kinv[cc + 3 * r];
                }
            }
        m3_mul(C1, F, CF);
        // C_1 is exactly symmetric (its off-diagonal entries are mu a + mu b in either
        // order): its six distinct entries 00 10 20 11 21 22 follow Q
        constexpr int SU[6] = {0, 1, 2, 4, 5, 8};
#pragma unroll
        for (int k = 0; k < 9; k++) {
            A.pstate[15 * p + k] = kill ? 0.0 : PK[k] - CF[k];
            if (store) A.defgrad[9 * p + k] = F[k];
        }
#pragma unroll
        for (int m = 0; m < 6; m++) A.pstate[15 * p + 9 + m] = kill ? 0.0 : C1[SU[m]];
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) {
            // a killed point's F may hold NaN (a singular K once all its bonds are off):
            // its state is zero, so no reader can pick the NaN up
            A.pstate[9 * p + k] = kill ? 0.0 : PK[k] - beta * F[k];
            if (store) A.defgrad[9 * p + k] = F[k];
        }
    }
    if (kill && !dead) A.killed[p] = 1;
    if (store) {
        A.nact_out[p] = kill ? 0 : na;
        A.dmg_out[p] = dmg;
    }
    return kill;
}

// CPD pass 1: calc_failure! (bond_system.jl:309-327), calc_damage!,
// calc_deformation_gradient! (correspondence.jl:193-214), too_much_damage! (:288-298),
// calc_first_piola_kirchhoff! (:216-224).  Stores Q = P K^-1 - beta F of every point: the
// ZEMSilling bond force of pass 2, w [(P K^-1) dX + beta (dx - F dX)] = w (Q dX + beta dx),
// then needs 9 numbers per neighbour instead of 18.
// S lanes per point take interleaved bonds; their sums are folded with lane shuffles and
// one lane finishes the point.
template <int S>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_cpd_pass1(pdx_step_args A)
{
    const int sub = threadIdx.x % S;
    const int64_t p = pdx_xcd_block() * PDX_BLOCK + threadIdx.x / S;
    if (p >= A.n) return;  // all S lanes of the point together
    const int c = A.cnt[p];
    const int64_t b0 = pdx_bond0(A.slice_ptr, p);
    const double4 ri = A.rec_in[p];
    const double4 Xi = A.Xr[p];
    const bool dead = A.killed[p] != 0;
    double Km[6] = {0, 0, 0, 0, 0, 0};  // 00 10 20 11 21 22
    double Ft[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int na = 0;
    constexpr int U = 2;  // bonds per lane in flight
    for (int t0 = sub; t0 < c; t0 += U * S) {
        uint32_t wdu[U];
        double Lu[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const bool ok = t0 + u * S < c;
            wdu[u] = ok ? A.bw[b0 + (int64_t)(t0 + u * S) * 64] : 0u;
            Lu[u] = ok ? A.L[b0 + (int64_t)(t0 + u * S) * 64] : 1.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (t0 + u * S >= c) continue;
            const int64_t b = b0 + (int64_t)(t0 + u * S) * 64;
            const uint32_t wd = wdu[u];
            const uint32_t j = wd & PDX_BW_NBR;
            const double4 rj = A.rec_in[j];
            const double4 Xj = A.Xr[j];
            const double dx[3] = {rj.x - ri.x, rj.y - ri.y, rj.z - ri.z};
            const double DX[3] = {Xj.x - Xi.x, Xj.y - Xi.y, Xj.z - Xi.z};
            const double Lb = Lu[u];
            double rl;
            const double l = pdx_sqrt_rsq((dx[0] * dx[0] + dx[1] * dx[1]) + dx[2] * dx[2], rl);
            const double e = l - Lb;
            const double rL = pdx_rcp1(Lb);
            const bool act = bond_fail(A, A.bw + b, wd, e, Lb, e * rL);
            if (act && !dead) {
                na++;
                const double temp = pdx_omega(A.kernel, A.delta, Lb, rL) * rj.w;
                Km[0] += temp * (DX[0] * DX[0]);
                Km[1] += temp * (DX[1] * DX[0]);
                Km[2] += temp * (DX[2] * DX[0]);
                Km[3] += temp * (DX[1] * DX[1]);
                Km[4] += temp * (DX[2] * DX[1]);
                Km[5] += temp * (DX[2] * DX[2]);
#pragma unroll
                for (int cc = 0; cc < 3; cc++)
#pragma unroll
                    for (int r = 0; r < 3; r++) Ft[r + 3 * cc] += temp * (dx[r] * DX[cc]);
            }
        }
    }
#pragma unroll
    for (int o = 1; o < S; o <<= 1) {
#pragma unroll
        for (int q = 0; q < 6; q++) Km[q] += __shfl_xor(Km[q], o);
#pragma unroll
        for (int q = 0; q < 9; q++) Ft[q] += __shfl_xor(Ft[q], o);
        na += __shfl_xor(na, o);
    }
    if (sub != 0) return;
    cpd_point_state(A, p, Km, Ft, na, c, dead, A.lambda, A.mu, A.G, A.K, A.nu, A.beta_zem);
}

// CPD pass 2: c_force_density! with ZEMSilling (correspondence.jl:226-251) as a gather:
// own term  t_ij V_j           with t_ij = w [(PKinv_i) dX + beta (dx - F_i dX)]
//                                        = w (Q_i dX + beta dx)
// reverse  -t_ji V_i  = V_i w (Q_j dX + beta dx)
template <int S>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_cpd_pass2(pdx_step_args A)
{
    const int sub = threadIdx.x % S;
    const int64_t p = pdx_xcd_block() * PDX_BLOCK + threadIdx.x / S;
    if (p >= A.n) return;  // all S lanes of the point together
    const int c = A.cnt[p];
    const int64_t b0 = pdx_bond0(A.slice_ptr, p);
    const double4 ri = A.rec_in[p];
    const double4 Xi = A.Xr[p];
    const bool dead_i = A.killed[p] != 0;
    const double beta = A.beta_zem;
    double Qi[9];
#pragma unroll
    for (int k = 0; k < 9; k++) Qi[k] = A.pstate[9 * p + k];
    double b3[3] = {0.0, 0.0, 0.0};
    constexpr int U = 2;  // bonds per lane in flight
    for (int t0 = sub; t0 < c; t0 += U * S) {
        uint32_t wdu[U];
        double Lu[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const bool ok = t0 + u * S < c;
            wdu[u] = ok ? A.bw[b0 + (int64_t)(t0 + u * S) * 64] : 0u;
            Lu[u] = ok ? A.L[b0 + (int64_t)(t0 + u * S) * 64] : 1.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t wd = wdu[u];
            if (!(wd & PDX_BW_ACT)) continue;  // also the slots past the row
            const uint32_t j = wd & PDX_BW_NBR;
            const double4 rj = A.rec_in[j];
            const double4 Xj = A.Xr[j];
            const double dx[3] = {rj.x - ri.x, rj.y - ri.y, rj.z - ri.z};
            const double DX[3] = {Xj.x - Xi.x, Xj.y - Xi.y, Xj.z - Xi.z};
            const double wv = pdx_omega(A.kernel, A.delta, Lu[u], pdx_rcp1(Lu[u])) * rj.w;  // w V_j
            const bool live_j = !A.killed[j];
            const double *qj = A.pstate + 9 * (int64_t)j;
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const double bdx = beta * dx[r];
                double tt = 0.0;
                if (!dead_i) tt += ((Qi[r] * DX[0] + Qi[r + 3] * DX[1]) + Qi[r + 6] * DX[2]) + bdx;
                if (live_j) tt += ((qj[r] * DX[0] + qj[r + 3] * DX[1]) + qj[r + 6] * DX[2]) + bdx;
                b3[r] += wv * tt;
            }
        }
    }
#pragma unroll
    for (int o = 1; o < S; o <<= 1)
#pragma unroll
        for (int r = 0; r < 3; r++) b3[r] += __shfl_xor(b3[r], o);
    static_assert(S == 4, "the tail takes four lanes per point");
    point_tail4(A, p, ri, b3, sub, A.rho);
}

// bonds per group of CPD pass 2's bond walk (pdx_walk_tile): the loads of the next group are
// in flight while a group is evaluated.  Measured at 512 k points, pass 2 with ZEMSilling:
// 0.330 / 0.306 / 0.319 / 0.328 / 0.365 ms for groups of 1 / 2 / 3 / 4 / 8 (ZEMWan within
// 1 % for 2 to 4).  Pass 1, whose bond arithmetic is longer, runs fastest with groups of 1.
#ifndef PDX_CPD_GROUP
#define PDX_CPD_GROUP 2
#endif

// ----------------------------------------------------------------------------------------
// CPD passes on the LDS-staged layout (FP64), for neighbour tables longer than one LDS
// budget: the block's table is staged in tiles of TL entries.  A point's bonds are sorted
// by neighbour index, as is the table, so each lane's bonds fall into the tiles in order:
// the lane walks its bonds while their slots lie in the staged tile and resumes there after
// the next tile is staged.  Pass 1 entries: x|V, X (64 B); pass 2: x|V, X|beta [j alive],
// [j alive] Q_j (136 B).  The arithmetic is that of k_cpd_pass1/2.
// ----------------------------------------------------------------------------------------
template <int S, int F>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_cpd_lds1(pdx_step_args A)
{
    // PDX_FEAT_MULTI: several parameter sets (force_density_point! with the params of the
    // point whose loop runs, correspondence.jl:183-191): eps_c of the owner per copy of a
    // bond, the neighbour's copy tracked in RACT as in the OSB pass; lambda, mu, G, K, beta
    // of the point
    constexpr bool MULTI = (F & PDX_FEAT_MULTI) != 0;
    extern __shared__ double4 hrec[];
    const int TL = pdx_lds_entries(A.halo_max, A.lds_budget, 64 + (MULTI ? 1 : 0));
    double4 *const hx = hrec;
    double4 *const hX = hrec + TL;
    uint8_t *const hpid = reinterpret_cast<uint8_t *>(hrec + 2 * TL);
    const int64_t k = pdx_xcd_block();
    const int64_t p0 = k * PDX_BLOCK;
    if (p0 >= A.n) return;  // whole workgroup
    const int64_t h0 = A.halo_ptr[k];
    const int H = (int)(A.halo_ptr[k + 1] - h0);
    const int32_t *const hidx = A.halo + h0;
    const int sub = threadIdx.x % S;
    const int64_t p = p0 + threadIdx.x / S;
    const bool valid = p < A.n;  // the block's threads all take part in the tile barriers
    const int c = valid ? A.cnt[p] : 0;
    const int64_t b0 = valid ? pdx_bond0(A.slice_ptr, p) : 0;
    uint16_t *const wbase = A.bw16 + b0;
    const double *const Lbase = A.L + b0;
    const double4 ri = valid ? A.rec_in[p] : make_double4(0.0, 0.0, 0.0, 0.0);
    const double4 Xi = valid ? A.Xr[p] : make_double4(0.0, 0.0, 0.0, 0.0);
    const bool dead = valid && A.killed[p] != 0;
    double epsc = A.epsc, epsc_lo = A.epsc_lo, epsc_hi = A.epsc_hi;
    double lam = A.lambda, mu = A.mu, Gp = A.G, Kp = A.K, nup = A.nu, beta = A.beta_zem;
    if (MULTI && valid) {
        const pdx_pset Pi = A.psets[A.pid[p]];
        epsc = Pi.epsc;
        epsc_lo = Pi.epsc_lo;
        epsc_hi = Pi.epsc_hi;
        lam = Pi.lam;
        mu = Pi.mu;
        Gp = Pi.G;
        Kp = Pi.K;
        nup = Pi.nu;
        beta = Pi.beta;
    }
    double Km[6] = {0, 0, 0, 0, 0, 0};  // 00 10 20 11 21 22
    double Ft[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    // the sums of w V_j DX and w V_j dx over the owner's copies of the bonds, for the own
    // bond terms of pass 2
    constexpr bool ZWF = (F & PDX_FEAT_ZW) != 0;
    double S1[3] = {0, 0, 0}, S2[3] = {0, 0, 0};
    int na = 0;
    int t = sub;
    for (int T0 = 0; T0 < H; T0 += TL) {
        const int T1 = H < T0 + TL ? H : T0 + TL;
        if (T0 > 0) __syncthreads();  // the previous tile is consumed
        for (int h = T0 + (int)threadIdx.x; h < T1; h += PDX_BLOCK * S) {
            const int32_t j = hidx[h];
            hx[h - T0] = A.rec_in[j];
            hX[h - T0] = A.Xr[j];
            if (MULTI) hpid[h - T0] = A.pid[j];
        }
        __syncthreads();
        t = pdx_walk_tile<1, S>(wbase, Lbase, t, c, T1, [&](uint32_t wd, double Lb, int tb) {
            const int slot = (int)(wd & PDX_W16_SLOT);
            const double4 rj = hx[slot - T0];
            const double4 Xj = hX[slot - T0];
            const double dx[3] = {rj.x - ri.x, rj.y - ri.y, rj.z - ri.z};
            const double DX[3] = {Xj.x - Xi.x, Xj.y - Xi.y, Xj.z - Xi.z};
            double rl;
            const double l = pdx_sqrt_rsq((dx[0] * dx[0] + dx[1] * dx[1]) + dx[2] * dx[2], rl);
            const double e = l - Lb;
            const double rL = pdx_rcp1(Lb);
            bool act;
            if (MULTI) {
                // each end fails its copy of the bond against its own eps_c; the stretch is the
                // same number at both ends
                const double eps = e * rL;
                const pdx_pset Pj = A.psets[hpid[slot - T0]];
                uint32_t nw = wd;
                if ((wd & PDX_W16_ACT) && (wd & PDX_W16_FP) && eps >= epsc_lo)
                    if ((eps > epsc_hi) ? true : ((e / Lb) > epsc)) nw &= ~PDX_W16_ACT;
                if ((wd & PDX_W16_RACT) && (wd & PDX_W16_FP) && eps >= Pj.epsc_lo)
                    if ((eps > Pj.epsc_hi) ? true : ((e / Lb) > Pj.epsc)) nw &= ~PDX_W16_RACT;
                if (nw != wd) wbase[(int64_t)tb * 64] = (uint16_t)nw;
                act = (nw & PDX_W16_ACT) != 0;
            } else {
                act = bond_fail16(wbase + (int64_t)tb * 64, wd, e, Lb, e * rL, epsc, epsc_lo, epsc_hi);
            }
            if (act && !dead) {
                na++;
                // K += w V_j DX DX', Ft += w V_j dx DX' as fused products of (w V_j DX)
                const double temp = pdx_omega(A.kernel, A.delta, Lb, rL) * rj.w;
                const double tD[3] = {temp * DX[0], temp * DX[1], temp * DX[2]};
                Km[0] = fma(tD[0], DX[0], Km[0]);
                Km[1] = fma(tD[1], DX[0], Km[1]);
                Km[2] = fma(tD[2], DX[0], Km[2]);
                Km[3] = fma(tD[1], DX[1], Km[3]);
                Km[4] = fma(tD[2], DX[1], Km[4]);
                Km[5] = fma(tD[2], DX[2], Km[5]);
#pragma unroll
                for (int cc = 0; cc < 3; cc++)
#pragma unroll
                    for (int r = 0; r < 3; r++) Ft[r + 3 * cc] = fma(tD[cc], dx[r], Ft[r + 3 * cc]);
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    S1[r] += tD[r];
                    S2[r] = fma(temp, dx[r], S2[r]);
                }
            }
        });
    }
#pragma unroll
    for (int o = 1; o < S; o <<= 1) {
#pragma unroll
        for (int q = 0; q < 6; q++) Km[q] += __shfl_xor(Km[q], o);
#pragma unroll
        for (int q = 0; q < 9; q++) Ft[q] += __shfl_xor(Ft[q], o);
#pragma unroll
        for (int r = 0; r < 3; r++) {
            S1[r] += __shfl_xor(S1[r], o);
            S2[r] += __shfl_xor(S2[r], o);
        }
        na += __shfl_xor(na, o);
    }
    if (!valid || sub != 0) return;
    const bool kill = cpd_point_state<ZWF>(A, p, Km, Ft, na, c, dead, lam, mu, Gp, Kp, nup, beta);
    {
        // the own term of the point's bonds, Q_i S1 + beta S2, and a_i = Q_i X_i + beta x_i
        // (ZEMWan: C_1i in place of beta; zero for a dead point, whose state is zero), behind
        // the state block (x_i, X_i read again rather than held through the point state:
        // its stress code takes the registers)
        constexpr int NQ = ZWF ? 15 : 9;
        constexpr int SY[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};  // (r, c) -> stored entry of C_1
        const double *const Q = A.pstate + NQ * p;
        const double *const C1 = Q + 9;
        const double be = kill ? 0.0 : beta;
        const double4 Xp = A.Xr[p], xp = A.rec_in[p];
        const double X3[3] = {Xp.x, Xp.y, Xp.z}, x3[3] = {xp.x, xp.y, xp.z};
        double *const ex = A.pstate + NQ * A.n + 6 * p;
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const double qs = fma(Q[r + 6], S1[2], fma(Q[r + 3], S1[1], Q[r] * S1[0]));
            const double qx = fma(Q[r + 6], X3[2], fma(Q[r + 3], X3[1], Q[r] * X3[0]));
            if (ZWF) {
                ex[r] = fma(C1[SY[r + 6]], S2[2], fma(C1[SY[r + 3]], S2[1], fma(C1[SY[r]], S2[0], qs)));
                ex[3 + r] = fma(C1[SY[r + 6]], x3[2], fma(C1[SY[r + 3]], x3[1], fma(C1[SY[r]], x3[0], qx)));
            } else {
                ex[r] = fma(be, S2[r], qs);
                ex[3 + r] = fma(be, x3[r], qx);
            }
        }
    }
}

// CPD pass 2 (correspondence.jl:234-249).  The own terms of the point's bonds, w V_j [Q_i DX +
// beta dx] summed, came from pass 1 as Q_i S1 + beta S2; the neighbour's term of a bond,
// Q_j DX + beta_j dx, is a_j - Q_j X_i - beta_j x_i with a_j = Q_j X_j + beta_j x_j from
// pass 1, in coordinates relative to the block's first point (the staging moves a_j there),
// so that the products stay of the size of the block rather than of the coordinates.  ZW
// (ZEMWan): the symmetric C_1 of each end in place of beta.  Table entries as 16-byte pairs:
// [j alive] (Q_j, a_j), then [j alive] beta, V_j (112 B, 1463 entries in a 160 KB tile), or
// with ZEMWan [j alive] C_1j (six entries), V_j and a pad (160 B, 1024 entries), against
// the 136 B (184 B) of x, X and the state in a per-bond form.  PDX_FEAT_MULTI (several
// parameter sets): beta_j of the neighbour's set, the neighbour's term for the neighbour's
// copy of the bond (RACT; the own terms followed the owner's copy, ACT, in pass 1), and the
// point's own rho (velocity_verlet.jl:262-275).
template <int S, int F>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_cpd_lds2_own(pdx_step_args A)
{
    constexpr bool MULTI = (F & PDX_FEAT_MULTI) != 0;
    constexpr bool ZW = (F & PDX_FEAT_ZW) != 0;
    constexpr uint32_t NBIT = MULTI ? PDX_W16_RACT : PDX_W16_ACT;  // the neighbour's copy
    constexpr int NQ = ZW ? 15 : 9;  // state numbers per point
    constexpr int ND = ZW ? 10 : 7;  // 16-byte pairs per entry
    constexpr int IV = ZW ? 18 : 13;  // V_j in the entry
    constexpr int SY[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};  // (r, c) -> stored entry of C_1
    extern __shared__ double4 hrec[];
    double2 *const hq = reinterpret_cast<double2 *>(hrec);
    const int TL = pdx_lds_entries(A.halo_max, A.lds_budget, 16 * ND);
    const int64_t k = pdx_xcd_block();
    const int64_t p0 = k * PDX_BLOCK;
    if (p0 >= A.n) return;  // whole workgroup
    const int64_t h0 = A.halo_ptr[k];
    const int H = (int)(A.halo_ptr[k + 1] - h0);
    const int32_t *const hidx = A.halo + h0;
    const int sub = threadIdx.x % S;
    const int64_t p = p0 + threadIdx.x / S;
    const bool valid = p < A.n;
    const int c = valid ? A.cnt[p] : 0;
    const int64_t b0 = valid ? pdx_bond0(A.slice_ptr, p) : 0;
    const uint16_t *const wbase = A.bw16 + b0;
    const double *const Lbase = A.L + b0;
    const double4 OX = A.Xr[p0], Ox = A.rec_in[p0];
    const double4 ri = valid ? A.rec_in[p] : Ox;
    const double4 Xi = valid ? A.Xr[p] : OX;
    const double xr[3] = {ri.x - Ox.x, ri.y - Ox.y, ri.z - Ox.z};
    const double Xr[3] = {Xi.x - OX.x, Xi.y - OX.y, Xi.z - OX.z};
    const double O3x[3] = {Ox.x, Ox.y, Ox.z}, O3X[3] = {OX.x, OX.y, OX.z};
    const double *const ext = A.pstate + NQ * A.n;
    const double rho = MULTI && valid ? A.psets[A.pid[p]].rho : A.rho;
    double b3[3] = {0.0, 0.0, 0.0};
    int t = sub;
    for (int T0 = 0; T0 < H; T0 += TL) {
        const int T1 = H < T0 + TL ? H : T0 + TL;
        if (T0 > 0) __syncthreads();  // the previous tile is consumed
        for (int h = T0 + (int)threadIdx.x; h < T1; h += PDX_BLOCK * S) {
            const int64_t j = hidx[h];
            const bool live = !A.killed[j];
            const double *const q = A.pstate + NQ * j;
            const double *const aj = ext + 6 * j + 3;
            double e[2 * ND];
#pragma unroll
            for (int m = 0; m < 9; m++) e[m] = live ? q[m] : 0.0;
            if (ZW) {
#pragma unroll
                for (int m = 0; m < 6; m++) e[12 + m] = live ? q[9 + m] : 0.0;
                e[19] = 0.0;
            } else {
                e[12] = live ? (MULTI ? A.psets[A.pid[j]].beta : A.beta_zem) : 0.0;
            }
            e[IV] = A.rec_in[j].w;
#pragma unroll
            for (int r = 0; r < 3; r++) {
                double o = fma(e[r + 6], O3X[2], fma(e[r + 3], O3X[1], e[r] * O3X[0]));
                if (ZW)
                    o = fma(e[12 + SY[r + 6]], O3x[2], fma(e[12 + SY[r + 3]], O3x[1], fma(e[12 + SY[r]], O3x[0], o)));
                else
                    o = fma(e[12], O3x[r], o);
                e[9 + r] = live ? aj[r] - o : 0.0;
            }
            double2 *const d = hq + ND * (int64_t)(h - T0);
#pragma unroll
            for (int m = 0; m < ND; m++) d[m] = make_double2(e[2 * m], e[2 * m + 1]);
        }
        __syncthreads();
        t = pdx_walk_tile<PDX_CPD_GROUP, S>(wbase, Lbase, t, c, T1, [&](uint32_t wd, double Lb, int) {
            if (!(wd & NBIT)) return;
            const double2 *const d = hq + ND * (int64_t)((int)(wd & PDX_W16_SLOT) - T0);
            double e[2 * ND];
#pragma unroll
            for (int m = 0; m < ND; m++) {
                const double2 v = d[m];
                e[2 * m] = v.x;
                e[2 * m + 1] = v.y;
            }
            const double wv = pdx_omega(A.kernel, A.delta, Lb, pdx_rcp1(Lb)) * e[IV];  // w V_j
#pragma unroll
            for (int r = 0; r < 3; r++) {
                double o = fma(e[r + 6], Xr[2], fma(e[r + 3], Xr[1], e[r] * Xr[0]));
                if (ZW)
                    o = fma(e[12 + SY[r + 6]], xr[2], fma(e[12 + SY[r + 3]], xr[1], fma(e[12 + SY[r]], xr[0], o)));
                else
                    o = fma(e[12], xr[r], o);
                b3[r] = fma(wv, e[9 + r] - o, b3[r]);
            }
        });
    }
#pragma unroll
    for (int o = 1; o < S; o <<= 1)
#pragma unroll
        for (int r = 0; r < 3; r++) b3[r] += __shfl_xor(b3[r], o);
    if (!valid) return;
#pragma unroll
    for (int r = 0; r < 3; r++) b3[r] += ext[6 * p + r];  // the own terms (pass 1)
    static_assert(S == 4, "the tail takes four lanes per point");
    point_tail4(A, p, ri, b3, sub, rho);
}

// ----------------------------------------------------------------------------------------
// Mixed-precision CPD (LDS-staged, as the mixed OSB passes).  The deformation gradient is
// formed as F = I + H with H = (sum w V_j U (x) D) K^-1, D the reference bond vector and U
// the displacement difference in FP32 (block-relative, as in k_bb_lds_mx); K and the H sum
// are accumulated in FP64 from FP32 bond terms, and K^-1, F, P and the per-point state are
// FP64.  The ZEMSilling bond force w [(P K^-1) D + beta (d - F D)] is written as
// w (Q' D + beta U) with Q' = P K^-1 - beta H: U - H D is the small difference d - F D
// without the cancellation of d - F D in FP32.  Pass 1 stores Q' (9 numbers per point);
// pass 2 stages (D, V), (U, beta [j alive]), [j alive] Q'_j (68 B per table entry) and
// gathers both terms of every bond.
// ----------------------------------------------------------------------------------------
template <int U, int S, int F>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_cpd_pass1_mx(pdx_step_args A)
{
    constexpr bool CAP = (F & PDX_FEAT_CAP) != 0;
    extern __shared__ float4 hmx[];
    const int64_t k = pdx_xcd_block();
    const int64_t p0 = k * PDX_BLOCK;
    if (p0 >= A.n) return;  // whole workgroup
    const int64_t h0 = A.halo_ptr[k];
    const int H = (int)(A.halo_ptr[k + 1] - h0);
    const double4 O = A.Xr[p0], xo = A.rec_in[p0];
    const double uo0 = xo.x - O.x, uo1 = xo.y - O.y, uo2 = xo.z - O.z;
    const int HL = CAP ? pdx_lds_entries(A.halo_max, A.lds_budget, 32) : A.halo_max;
    const int Hs = H < HL ? H : HL;
    const int32_t *const hidx = A.halo + h0;
    for (int h = threadIdx.x; h < Hs; h += PDX_BLOCK * S) {
        float4 a, b;
        mx_stage_xu(A, hidx[h], O, uo0, uo1, uo2, a, b);
        hmx[2 * h] = a;
        hmx[2 * h + 1] = b;
    }
    __syncthreads();
    const int sub = threadIdx.x % S;
    const int64_t p = p0 + threadIdx.x / S;
    if (p >= A.n) return;  // all S lanes of the point together
    const int c = A.cnt[p];
    const int64_t b0 = pdx_bond0(A.slice_ptr, p);
    uint16_t *const wbase = A.bw16 + b0;
    const double4 ri = A.rec_in[p];
    const double4 Xi = A.Xr[p];
    const float xi = (float)(Xi.x - O.x), yi = (float)(Xi.y - O.y), zi = (float)(Xi.z - O.z);
    const float ui = (float)((ri.x - Xi.x) - uo0), vi = (float)((ri.y - Xi.y) - uo1),
                wi = (float)((ri.z - Xi.z) - uo2);
    const bool dead = A.killed[p] != 0;
    double Km[6] = {0, 0, 0, 0, 0, 0};  // 00 10 20 11 21 22
    double Ht[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int na = 0;
    const double epsc = A.epsc;
    const float epscf = (float)epsc, deltaf = (float)A.delta;
    const float win_lo = epscf * 0.9999f, win_hi = epscf * 1.0001f;
    for (int t = sub; t < c; t += U * S) {
        uint32_t wd[U];
#pragma unroll
        for (int u = 0; u < U; u++) wd[u] = (t + u * S < c) ? wbase[(int64_t)(t + u * S) * 64] : 0u;
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (t + u * S >= c) continue;
            const int slot = (int)(wd[u] & PDX_W16_SLOT);
            float4 a, b;
            if (!CAP || slot < HL) {
                a = hmx[2 * slot];
                b = hmx[2 * slot + 1];
            } else {
                mx_stage_xu(A, hidx[slot], O, uo0, uo1, uo2, a, b);
            }
            const float D[3] = {a.x - xi, a.y - yi, a.z - zi};
            const float Uv[3] = {b.x - ui, b.y - vi, b.z - wi};
            const float L2 = __builtin_fmaf(D[2], D[2], __builtin_fmaf(D[1], D[1], D[0] * D[0]));
            const float DU = __builtin_fmaf(D[2], Uv[2], __builtin_fmaf(D[1], Uv[1], D[0] * Uv[0]));
            const float UU = __builtin_fmaf(Uv[2], Uv[2], __builtin_fmaf(Uv[1], Uv[1], Uv[0] * Uv[0]));
            const float num = __builtin_fmaf(2.f, DU, UU);  // l^2 - L^2
            const float Lb = __builtin_amdgcn_sqrtf(L2);
            const float l = __builtin_amdgcn_sqrtf(L2 + num);
            const float rL = __builtin_amdgcn_rcpf(Lb);
            const float eps = (num * __builtin_amdgcn_rcpf(l + Lb)) * rL;
            bool act = (wd[u] & PDX_W16_ACT) != 0;
            if (act && (wd[u] & PDX_W16_FP) && eps > win_lo) {
                bool brk = eps > win_hi;
                if (!brk) {
                    const double L64 = A.L[b0 + (int64_t)(t + u * S) * 64];
                    const double4 xj = A.rec_in[hidx[slot]];
                    const double dx = xj.x - ri.x, dy = xj.y - ri.y, dz = xj.z - ri.z;
                    double rl;
                    const double l64 = pdx_sqrt_rsq((dx * dx + dy * dy) + dz * dz, rl);
                    brk = ((l64 - L64) / L64) > epsc;
                }
                if (brk) {
                    act = false;
                    wbase[(int64_t)(t + u * S) * 64] = (uint16_t)(wd[u] & ~PDX_W16_ACT);
                }
            }
            if (act && !dead) {
                na++;
                const float wv = (deltaf * rL) * a.w;  // w V_j
                const float wD[3] = {wv * D[0], wv * D[1], wv * D[2]};  // the products of w V_j D
                Km[0] += (double)(wD[0] * D[0]);
                Km[1] += (double)(wD[1] * D[0]);
                Km[2] += (double)(wD[2] * D[0]);
                Km[3] += (double)(wD[1] * D[1]);
                Km[4] += (double)(wD[2] * D[1]);
                Km[5] += (double)(wD[2] * D[2]);
#pragma unroll
                for (int cc = 0; cc < 3; cc++)
#pragma unroll
                    for (int r = 0; r < 3; r++) Ht[r + 3 * cc] += (double)(Uv[r] * wD[cc]);
            }
        }
    }
#pragma unroll
    for (int o = 1; o < S; o <<= 1) {
#pragma unroll
        for (int q = 0; q < 6; q++) Km[q] += __shfl_xor(Km[q], o);
#pragma unroll
        for (int q = 0; q < 9; q++) Ht[q] += __shfl_xor(Ht[q], o);
        na += __shfl_xor(na, o);
    }
    if (sub != 0) return;
    const double K9[9] = {Km[0], Km[1], Km[2], Km[1], Km[3], Km[4], Km[2], Km[4], Km[5]};
    double Kinv[9], Hm[9], Fm[9];
    m3_inv(K9, Kinv);
    m3_mul(Ht, Kinv, Hm);
#pragma unroll
    for (int q = 0; q < 9; q++) Fm[q] = Hm[q] + ((q % 4 == 0) ? 1.0 : 0.0);
    const double dmg = 1.0 - (double)na / (double)c;
    bool hasnan = false;
#pragma unroll
    for (int q = 0; q < 9; q++) hasnan |= isnan(Fm[q]);
    const bool kill = dead || dmg > A.maxdmg || hasnan;
    const bool store = (A.flags & (PDX_F_STORE | PDX_F_FORCE_ONLY)) != 0;
    double PK[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (!kill) {
        double P[9];
        pk1(A.model, Fm, A.lambda, A.mu, A.G, A.K, A.nu, P);
        m3_mul(P, Kinv, PK);
        if (store) {
            const double sj = 1.0 / m3_det(Fm);
#pragma unroll
            for (int q = 0; q < 9; q++) {
                const int r = q % 3, cc = q / 3;
                A.cauchy[9 * p + q] =
                    ((sj * P[r]) * Fm[cc] + (sj * P[r + 3]) * Fm[cc + 3]) + (sj * P[r + 6]) * Fm[cc + 6];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 9; q++) {
        A.pstate[9 * p + q] = kill ? 0.0 : PK[q] - A.beta_zem * Hm[q];  // no NaN of a dead K
        if (store) A.defgrad[9 * p + q] = Fm[q];
    }
    if (kill && !dead) A.killed[p] = 1;
    if (store) {
        A.nact_out[p] = kill ? 0 : na;
        A.dmg_out[p] = dmg;
    }
}

// the staged values of neighbour j for pass 2: (D, V), (U, beta [j alive]), [j alive] Q'_j
__device__ __forceinline__ void cpd_mx_entry(const pdx_step_args &A, int64_t j, double4 O, double uo0,
                                             double uo1, double uo2, float4 &a, float4 &b, float q[9])
{
    mx_stage_xu(A, j, O, uo0, uo1, uo2, a, b);
    const bool live = !A.killed[j];
    b.w = live ? (float)A.beta_zem : 0.f;
    // selects, not products with 0: a dead point's state must not leak NaN or Inf
#pragma unroll
    for (int r = 0; r < 9; r++) q[r] = live ? (float)A.pstate[9 * j + r] : 0.f;
}

template <int U, int S, int F>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_cpd_pass2_mx(pdx_step_args A)
{
    constexpr bool CAP = (F & PDX_FEAT_CAP) != 0;
    extern __shared__ float4 hmx[];  // [4 HL]: (D, V), (U, beta l), q0-3, q4-7; then q8 [HL]
    const int64_t k = pdx_xcd_block();
    const int64_t p0 = k * PDX_BLOCK;
    if (p0 >= A.n) return;  // whole workgroup
    const int64_t h0 = A.halo_ptr[k];
    const int H = (int)(A.halo_ptr[k + 1] - h0);
    const double4 O = A.Xr[p0], xo = A.rec_in[p0];
    const double uo0 = xo.x - O.x, uo1 = xo.y - O.y, uo2 = xo.z - O.z;
    const int HL = CAP ? pdx_lds_entries(A.halo_max, A.lds_budget, 68) : A.halo_max;
    const int Hs = H < HL ? H : HL;
    const int32_t *const hidx = A.halo + h0;
    float *const hq8 = reinterpret_cast<float *>(hmx + 4 * HL);
    for (int h = threadIdx.x; h < Hs; h += PDX_BLOCK * S) {
        float4 a, b;
        float q[9];
        cpd_mx_entry(A, hidx[h], O, uo0, uo1, uo2, a, b, q);
        hmx[4 * h] = a;
        hmx[4 * h + 1] = b;
        hmx[4 * h + 2] = make_float4(q[0], q[1], q[2], q[3]);
        hmx[4 * h + 3] = make_float4(q[4], q[5], q[6], q[7]);
        hq8[h] = q[8];
    }
    __syncthreads();
    const int sub = threadIdx.x % S;
    const int64_t p = p0 + threadIdx.x / S;
    if (p >= A.n) return;  // all S lanes of the point together
    const int c = A.cnt[p];
    const int64_t b0 = pdx_bond0(A.slice_ptr, p);
    const uint16_t *const wbase = A.bw16 + b0;
    const double4 ri = A.rec_in[p];
    const double4 Xi = A.Xr[p];
    const float xi = (float)(Xi.x - O.x), yi = (float)(Xi.y - O.y), zi = (float)(Xi.z - O.z);
    const float ui = (float)((ri.x - Xi.x) - uo0), vi = (float)((ri.y - Xi.y) - uo1),
                wi = (float)((ri.z - Xi.z) - uo2);
    const bool live_i = !A.killed[p];
    const float beta_i = live_i ? (float)A.beta_zem : 0.f;
    float Qi[9];
#pragma unroll
    for (int r = 0; r < 9; r++) Qi[r] = live_i ? (float)A.pstate[9 * p + r] : 0.f;
    const float deltaf = (float)A.delta;
    double b3[3] = {0.0, 0.0, 0.0};
    for (int t = sub; t < c; t += U * S) {
        uint32_t wd[U];
#pragma unroll
        for (int u = 0; u < U; u++) wd[u] = (t + u * S < c) ? wbase[(int64_t)(t + u * S) * 64] : 0u;
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (t + u * S >= c) continue;
            if (!(wd[u] & PDX_W16_ACT)) continue;
            const int slot = (int)(wd[u] & PDX_W16_SLOT);
            float4 a, b;
            float q[9];
            if (!CAP || slot < HL) {
                a = hmx[4 * slot];
                b = hmx[4 * slot + 1];
                const float4 q0 = hmx[4 * slot + 2], q1 = hmx[4 * slot + 3];
                q[0] = q0.x; q[1] = q0.y; q[2] = q0.z; q[3] = q0.w;
                q[4] = q1.x; q[5] = q1.y; q[6] = q1.z; q[7] = q1.w;
                q[8] = hq8[slot];
            } else {
                cpd_mx_entry(A, hidx[slot], O, uo0, uo1, uo2, a, b, q);
            }
            const float D[3] = {a.x - xi, a.y - yi, a.z - zi};
            const float Uv[3] = {b.x - ui, b.y - vi, b.z - wi};
            const float L2 = __builtin_fmaf(D[2], D[2], __builtin_fmaf(D[1], D[1], D[0] * D[0]));
            const float wv = (deltaf * __builtin_amdgcn_rsqf(L2)) * a.w;  // w V_j
#pragma unroll
            for (int r = 0; r < 3; r++) {
                // own term Q'_i D + beta U, reverse term Q'_j D + beta U (zero for dead ends)
                const float qs0 = Qi[r] + q[r], qs1 = Qi[r + 3] + q[r + 3], qs2 = Qi[r + 6] + q[r + 6];
                const float tt = __builtin_fmaf(qs2, D[2], __builtin_fmaf(qs1, D[1], qs0 * D[0])) +
                                 (beta_i + b.w) * Uv[r];
                b3[r] += (double)(wv * tt);
            }
        }
    }
#pragma unroll
    for (int o = 1; o < S; o <<= 1)
#pragma unroll
        for (int r = 0; r < 3; r++) b3[r] += __shfl_xor(b3[r], o);
    static_assert(S == 4, "the tail takes four lanes per point");
    point_tail4(A, p, ri, b3, sub, A.rho);
}

// ----------------------------------------------------------------------------------------
// prologue: the first half step of a run (update_vel_half!, BCs, update_disp_and_pos!)
// ----------------------------------------------------------------------------------------
__global__ void k_vv_prologue(int64_t n, const double *v, const double *a, double *vh,
                              double *u, double4 *rec, const int32_t *bcslot,
                              const int32_t *slot_src, const double *bcv, double dt, double dth)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * n) return;
    const int64_t i = t / 3;
    const int d = (int)(t - 3 * i);
    double vn = v[t] + a[t] * dth;
    const int32_t slot = bcslot[i];
    if (slot >= 0) {
        const int32_t s = slot_src[3 * slot + d];
        if (s >= 0) {
            const double val = bcv[s];
            if (isfinite(val)) vn = val;
        }
    }
    const double du = vn * dt;
    u[t] = u[t] + du;
    double *x = reinterpret_cast<double *>(rec) + 4 * i + d;
    *x = *x + du;
    vh[t] = vn;
}

// ----------------------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------------------
pdx_status pdx_fill_step_args(pdx_ctx *c, pdx_step_args *a, int64_t step, int flags, double dt,
                              const double *bcv_dev)
{
    std::memset(a, 0, sizeof *a);
    a->n = c->n;
    a->step = step;
    a->flags = flags | (c->bext_nonzero ? PDX_F_BEXT : 0);
    a->rho = c->P.rho;
    a->dt = dt;
    a->dth = 0.5 * dt;
    a->slice_ptr = c->slice_ptr;
    a->cnt = c->cnt;
    a->bw = c->bw;
    a->bw16 = c->bw16;
    a->halo_ptr = c->halo_ptr;
    a->halo = c->halo;
    a->halo_max = c->halo_max;
    a->lds_budget = pdx_lds_budget(c);
    a->psets = c->multi() ? c->d_psets : nullptr;
    a->pid = c->multi() ? c->d_pid : nullptr;
    a->scf = c->corr ? c->scf : nullptr;
    a->scf32 = c->corr ? c->scf32 : nullptr;
    a->L = c->L;
    a->Lf = c->Lf;
    a->rec_in = c->rec[c->cur];
    a->rec_out = c->rec[c->cur ^ 1];
    a->Xr = c->Xr;
    a->u = c->u;
    a->vh = c->vh;
    a->bext = c->bext;
    a->v_out = c->v;
    a->a_out = c->a;
    a->bint_out = c->bint;
    a->dmg_out = c->damage;
    a->nact_out = c->nact;
    a->bcslot = c->bcslot;
    a->slot_src = c->slot_src;
    a->bcv = bcv_dev;
    a->flags_out = c->d_flags;
    a->bc = c->P.bc;
    a->epsc = c->P.epsc;
    // (l - L) * (1/L) with a one-step reciprocal is within ~1e-13 of the correctly rounded
    // (l - L) / L; outside [epsc_lo, epsc_hi] the two compare equal against eps_c, inside
    // the kernel takes the correctly rounded quotient
    a->epsc_lo = c->P.epsc - 1e-11 * fabs(c->P.epsc);
    a->epsc_hi = c->P.epsc + 1e-11 * fabs(c->P.epsc);
    a->delta = c->P.delta;
    a->G = c->P.G;
    a->K = c->P.K;
    a->lambda = c->P.lambda;
    a->mu = c->P.mu;
    a->nu = c->P.nu;
    a->beta_zem = (c->P.zem_cs * c->P.bc) / c->P.delta;  // correspondence.jl:232
    a->maxdmg = c->P.maxdmg;
    a->model = c->P.model;
    a->kernel = c->kind == PDX_BB ? PDX_KERNEL_LINEAR : c->P.kernel;
    a->pstate = c->pstate;
    a->killed = c->killed;
    a->defgrad = c->defgrad;
    a->cauchy = c->cauchy;
    return PDX_OK;
}


// launches of the LDS-staged kernels with a feature mask F (PDX_FEAT_*)
static int pdx_features(const pdx_ctx *c)
{
    return (c->multi() ? PDX_FEAT_MULTI : 0) | (c->corr ? PDX_FEAT_SC : 0) |
           (c->kind == PDX_OSB && c->P.kernel != PDX_KERNEL_LINEAR ? PDX_FEAT_KERN : 0);
}

// A kernel whose table entries take `entry` LDS bytes stages the whole table if the longest
// one fits the budget; otherwise (PDX_FEAT_CAP) it stages what fits and reads the rest from
// global memory, so that two blocks still share a CU.  The capped variant is a separate
// instantiation: its fallback branch costs the uncapped loop ~30 %.
static bool lds_capped(const pdx_ctx *c, int entry, int budget)
{
    return (int64_t)std::max(c->halo_max, 1) * entry > budget;
}

static bool lds_capped(const pdx_ctx *c, int entry) { return lds_capped(c, entry, pdx_lds_budget(c)); }

static size_t lds_bytes(const pdx_ctx *c, int entry, int budget)
{
    const int hl = lds_capped(c, entry, budget) ? pdx_lds_entries(c->halo_max, budget, entry)
                                                : std::max(c->halo_max, 1);
    return (size_t)hl * entry + 16;
}

static size_t lds_bytes(const pdx_ctx *c, int entry) { return lds_bytes(c, entry, pdx_lds_budget(c)); }

// VGPRs per lane of a kernel (queried once): a 1024-thread block fits twice on a CU only
// with at most 64
static int kernel_vgprs(const void *fn)
{
    static std::mutex m;
    static std::unordered_map<const void *, int> cache;
    std::lock_guard<std::mutex> lock(m);
    auto it = cache.find(fn);
    if (it != cache.end()) return it->second;
    hipFuncAttributes at{};
    const int v = hipFuncGetAttributes(&at, fn) == hipSuccess ? at.numRegs : 0;
    cache[fn] = v;
    return v;
}

// Launch of an LDS-staged kernel whose table entries take `entry` bytes.  A capped variant
// (tables longer than the budget) whose registers allow one block per CU anyway may take the
// CU's whole LDS.  Measured at 576 k points: a tiled kernel gains from fewer tiles whenever
// the table is longer than the budget; a kernel that reads the entries past the budget from
// global memory gains once they are many (horizon 5.015 dx: ~80 % past the budget) and
// loses a little when they are few (4.015 dx: ~15 %).
static void launch_lds(pdx_ctx *c, const void *fn, bool capped, bool tiled, int entry, dim3 G,
                       dim3 BL, const pdx_step_args &a)
{
    pdx_step_args af = a;
    int bud = pdx_lds_budget(c);
    const bool many_past = (int64_t)c->halo_max * entry > (int64_t)bud * 3 / 2;
    if (capped && (tiled || many_past) && kernel_vgprs(fn) > 64) bud = pdx_lds_budget_full(c);
    af.lds_budget = bud;
    void *args[] = {&af};
    (void)hipLaunchKernel(fn, G, BL, args, lds_bytes(c, entry, bud), c->stream);
}

template <int F>
static void launch_bb_feat(pdx_ctx *c, bool mixed, dim3 G, dim3 BL, size_t lds, const pdx_step_args &a)
{
    if (mixed && (F & PDX_FEAT_CAP))
        launch_lds(c, (const void *)&k_bb_lds_mx<4, PDX_LDS_LANES, F>, true, false,
                   32 + ((F & PDX_FEAT_MULTI) ? 1 : 0), G, BL, a);
    else if (mixed) hipLaunchKernelGGL((k_bb_lds_mx<4, PDX_LDS_LANES, F>), G, BL, lds, c->stream, a);
    else hipLaunchKernelGGL((k_bb_lds<2, PDX_LDS_LANES, F>), G, BL, lds, c->stream, a);
}

// feature masks 1..7 (PDX_FEAT_MULTI | PDX_FEAT_SC | PDX_FEAT_CAP); for OSB also those with
// PDX_FEAT_KERN
#define PDX_FEAT_SWITCH(feat, CALL)                                                         \
    switch (feat) {                                                                         \
    case 1: CALL(1); break;                                                                 \
    case 2: CALL(2); break;                                                                 \
    case 3: CALL(3); break;                                                                 \
    case 4: CALL(4); break;                                                                 \
    case 5: CALL(5); break;                                                                 \
    case 6: CALL(6); break;                                                                 \
    case 7: CALL(7); break;                                                                 \
    default: break;                                                                         \
    }
#define PDX_FEAT_SWITCH_OSB(feat, CALL)                                                     \
    switch (feat) {                                                                         \
    case 16: CALL(16); break;                                                               \
    case 17: CALL(17); break;                                                               \
    case 18: CALL(18); break;                                                               \
    case 19: CALL(19); break;                                                               \
    case 20: CALL(20); break;                                                               \
    case 21: CALL(21); break;                                                               \
    case 22: CALL(22); break;                                                               \
    case 23: CALL(23); break;                                                               \
    default: PDX_FEAT_SWITCH(feat, CALL)                                                    \
    }

template <int F>
static void launch_osb_feat(pdx_ctx *c, int pass, dim3 G, dim3 BL, int entry, const pdx_step_args &a)
{
    const void *fn = pass == 1 ? (const void *)&k_osb_lds1<PDX_LDS_LANES, F>
                               : (const void *)&k_osb_lds2<PDX_LDS_LANES, F>;
    launch_lds(c, fn, (F & PDX_FEAT_CAP) != 0, true, entry, G, BL, a);
}

bool pdx_bb_jv_fusable(const pdx_ctx *c)
{
    return c->kind == PDX_BB && c->lds_mode && pdx_features(c) == 0 && !lds_capped(c, 32);
}

pdx_status pdx_launch_bb_jv(pdx_ctx *c, const double *u, const double *v, const double *un2,
                            const double *parts, int np, double sc, const uint8_t *cmask,
                            double *jv)
{
    if (!pdx_bb_jv_fusable(c)) return pdx_fail(c, PDX_ERR_UNSUPPORTED, "fused J v: layout not supported");
    pdx_step_args a;
    PDX_TRY(pdx_fill_step_args(c, &a, 0, PDX_F_FORCE_ONLY | PDX_F_SCALE_VOL | PDX_F_FP64, 0.0, nullptr));
    const int64_t blocks = (c->n + PDX_BLOCK - 1) / PDX_BLOCK;
    const dim3 G((unsigned)((blocks + PDX_NUM_XCD - 1) / PDX_NUM_XCD * PDX_NUM_XCD));
    const dim3 BL(PDX_BLOCK * PDX_LDS_LANES);
    const size_t two = (size_t)std::max(c->halo_max, 1) * 64 + 16;
    const char *env = std::getenv("PDX_JV_ONE_PASS");
    // (the static LDS of the sum adds 128 bytes)
    const bool one_pass = (int64_t)two + 256 <= c->lds_limit && !(env && env[0] == '0');
    if (one_pass)
        hipLaunchKernelGGL((k_bb_jv1<4, PDX_LDS_LANES>), G, BL, two, c->stream, a, u, v, un2, parts, np, sc,
                           cmask, jv);
    else
        hipLaunchKernelGGL((k_bb_jv<2, PDX_LDS_LANES>), G, BL, lds_bytes(c, 32), c->stream, a, u, v, un2,
                           parts, np, sc, cmask, jv);
    PDX_HIP(c, hipGetLastError());
    return PDX_OK;
}

pdx_status pdx_launch_step(pdx_ctx *c, const pdx_step_args &a)
{
    // the mixed-precision kernels, unless the caller needs the FP64 force (PDX_F_FP64)
    const bool mixed = c->precision == PDX_MIXED && !(a.flags & PDX_F_FP64);
    const int64_t blocks = (c->n + PDX_BLOCK - 1) / PDX_BLOCK;
    const dim3 G((unsigned)((blocks + PDX_NUM_XCD - 1) / PDX_NUM_XCD * PDX_NUM_XCD)), B(PDX_BLOCK);
    switch (c->kind) {
    case PDX_BB:
        PDX_TRY(pdx_timed_begin(c));
        if (c->lds_mode) {
            const size_t lds = lds_bytes(c, 32);
            // four lanes per point (the halo table of a block limits the blocks per CU, so
            // the lanes, not the blocks, supply the parallelism that hides the bond loads)
            const dim3 BL(PDX_BLOCK * PDX_LDS_LANES);
            // bonds per lane and unrolled iteration: the FP32 bond arithmetic of the mixed
            // step is short enough that four loads in flight pay; FP64 is indifferent
            const bool u4 = c->bb_unroll ? c->bb_unroll == 4 : mixed;
            // with several parameter sets each neighbour-table entry also carries the
            // neighbour's set
            const int entry = 32 + (c->multi() ? 1 : 0);
            const int feat = pdx_features(c) | (lds_capped(c, entry) ? PDX_FEAT_CAP : 0);
            const size_t ldsf = lds_bytes(c, entry);
#define PDX_BB_LAUNCH(F) launch_bb_feat<F>(c, mixed, G, BL, ldsf, a)
            if (feat) {
                PDX_FEAT_SWITCH(feat, PDX_BB_LAUNCH)
            } else if (mixed) {
                if (u4) hipLaunchKernelGGL((k_bb_lds_mx<4, PDX_LDS_LANES, 0>), G, BL, lds, c->stream, a);
                else hipLaunchKernelGGL((k_bb_lds_mx<2, PDX_LDS_LANES, 0>), G, BL, lds, c->stream, a);
            } else {
                if (u4) hipLaunchKernelGGL((k_bb_lds<4, PDX_LDS_LANES, 0>), G, BL, lds, c->stream, a);
                else hipLaunchKernelGGL((k_bb_lds<2, PDX_LDS_LANES, 0>), G, BL, lds, c->stream, a);
            }
        } else if (mixed) {
            if (c->bb_unroll == 4) hipLaunchKernelGGL((k_bb_step<true, 4>), G, B, 0, c->stream, a);
            else hipLaunchKernelGGL((k_bb_step<true, 2>), G, B, 0, c->stream, a);
        } else {
            if (c->bb_unroll == 4) hipLaunchKernelGGL((k_bb_step<false, 4>), G, B, 0, c->stream, a);
            else hipLaunchKernelGGL((k_bb_step<false, 2>), G, B, 0, c->stream, a);
        }
        PDX_HIP(c, hipGetLastError());
        PDX_TRY(pdx_timed_end(c, "step"));
        break;
    case PDX_OSB:
        // (an influence function other than linear_kernel runs the FP64 passes)
        if (c->lds_mode && mixed && !c->multi() && c->P.kernel == PDX_KERNEL_LINEAR) {
            const dim3 BL(PDX_BLOCK * PDX_LDS_LANES);
            // pass 1 entries (D, V), (U, 0): 32 B; pass 2 also (c2 V)_j: 36 B
            const int sc = c->corr ? PDX_FEAT_SC : 0;
            const int feat = sc | (lds_capped(c, 32) ? PDX_FEAT_CAP : 0);
            const int feat2 = sc | (lds_capped(c, 36) ? PDX_FEAT_CAP : 0);
            const size_t lds = lds_bytes(c, 32), lds2 = lds_bytes(c, 36);
            PDX_TRY(pdx_timed_begin(c));
            switch (feat) {
            case 0: hipLaunchKernelGGL((k_osb_lds1_mx<4, PDX_LDS_LANES, 0>), G, BL, lds, c->stream, a); break;
            case PDX_FEAT_SC: hipLaunchKernelGGL((k_osb_lds1_mx<4, PDX_LDS_LANES, PDX_FEAT_SC>), G, BL, lds, c->stream, a); break;
            case PDX_FEAT_CAP: launch_lds(c, (const void *)&k_osb_lds1_mx<4, PDX_LDS_LANES, PDX_FEAT_CAP>, true, false, 32, G, BL, a); break;
            default: launch_lds(c, (const void *)&k_osb_lds1_mx<4, PDX_LDS_LANES, PDX_FEAT_SC | PDX_FEAT_CAP>, true, false, 32, G, BL, a); break;
            }
            PDX_HIP(c, hipGetLastError());
            PDX_TRY(pdx_timed_end(c, "pass1"));
            PDX_TRY(pdx_timed_begin(c));
            switch (feat2) {
            case 0: hipLaunchKernelGGL((k_osb_lds2_mx<4, PDX_LDS_LANES, 0>), G, BL, lds2, c->stream, a); break;
            case PDX_FEAT_SC: hipLaunchKernelGGL((k_osb_lds2_mx<4, PDX_LDS_LANES, PDX_FEAT_SC>), G, BL, lds2, c->stream, a); break;
            case PDX_FEAT_CAP: launch_lds(c, (const void *)&k_osb_lds2_mx<4, PDX_LDS_LANES, PDX_FEAT_CAP>, true, false, 36, G, BL, a); break;
            default: launch_lds(c, (const void *)&k_osb_lds2_mx<4, PDX_LDS_LANES, PDX_FEAT_SC | PDX_FEAT_CAP>, true, false, 36, G, BL, a); break;
            }
            PDX_HIP(c, hipGetLastError());
            PDX_TRY(pdx_timed_end(c, "step"));
            break;
        }
        if (c->lds_mode) {
            const dim3 BL(PDX_BLOCK * PDX_LDS_LANES);
            const int mb = c->multi() ? 1 : 0;  // the neighbours' parameter sets
            // pass 2 entries: x and (c1 V, c2 V) with one set, x|V, (wvol, dil) and the set
            const int e2 = mb ? 48 + mb : 40;
            const int e1 = 32 + mb;
            const int f1 = pdx_features(c) | (lds_capped(c, e1) ? PDX_FEAT_CAP : 0);
            const int f2 = pdx_features(c) | (lds_capped(c, e2) ? PDX_FEAT_CAP : 0);
#define PDX_OSB1_LAUNCH(F) launch_osb_feat<F>(c, 1, G, BL, e1, a)
#define PDX_OSB2_LAUNCH(F) launch_osb_feat<F>(c, 2, G, BL, e2, a)
            PDX_TRY(pdx_timed_begin(c));
            if (f1) {
                PDX_FEAT_SWITCH_OSB(f1, PDX_OSB1_LAUNCH)
            } else {
                launch_osb_feat<0>(c, 1, G, BL, e1, a);
            }
            PDX_HIP(c, hipGetLastError());
            PDX_TRY(pdx_timed_end(c, "pass1"));
            PDX_TRY(pdx_timed_begin(c));
            if (f2) {
                PDX_FEAT_SWITCH_OSB(f2, PDX_OSB2_LAUNCH)
            } else {
                launch_osb_feat<0>(c, 2, G, BL, e2, a);
            }
            PDX_HIP(c, hipGetLastError());
            PDX_TRY(pdx_timed_end(c, "step"));
            break;
        }
        PDX_TRY(pdx_timed_begin(c));
        hipLaunchKernelGGL(k_osb_pass1, G, B, 0, c->stream, a);
        PDX_HIP(c, hipGetLastError());
        PDX_TRY(pdx_timed_end(c, "pass1"));
        PDX_TRY(pdx_timed_begin(c));
        hipLaunchKernelGGL(k_osb_pass2, G, B, 0, c->stream, a);
        PDX_HIP(c, hipGetLastError());
        PDX_TRY(pdx_timed_end(c, "step"));
        break;
    case PDX_CPD:
        if (c->lds_mode) {
            // The CPD passes hold 97-122 VGPRs, so one 1024-thread block runs per CU: their
            // tables take the CU's whole LDS (fewer tiles, fewer entries read past the cap).
            pdx_step_args af = a;
            const int bud = pdx_lds_budget_full(c);
            af.lds_budget = bud;
            const dim3 BL(PDX_BLOCK * PDX_LDS_LANES);
            if (mixed && !c->multi() && c->P.zem != PDX_ZEM_WAN &&
                c->P.kernel == PDX_KERNEL_LINEAR) {  // otherwise the FP64 passes
                // pass 1 entries (D, V), (U, 0): 32 B; pass 2 (D, V), (U, beta l), l Q': 68 B
                const size_t l1 = lds_bytes(c, 32, bud), l2b = lds_bytes(c, 68, bud);
                PDX_TRY(pdx_timed_begin(c));
                if (lds_capped(c, 32, bud))
                    hipLaunchKernelGGL((k_cpd_pass1_mx<1, PDX_LDS_LANES, PDX_FEAT_CAP>), G, BL, l1, c->stream, af);
                else
                    hipLaunchKernelGGL((k_cpd_pass1_mx<1, PDX_LDS_LANES, 0>), G, BL, l1, c->stream, af);
                PDX_HIP(c, hipGetLastError());
                PDX_TRY(pdx_timed_end(c, "pass1"));
                PDX_TRY(pdx_timed_begin(c));
                if (lds_capped(c, 68, bud))
                    hipLaunchKernelGGL((k_cpd_pass2_mx<2, PDX_LDS_LANES, PDX_FEAT_CAP>), G, BL, l2b, c->stream, af);
                else
                    hipLaunchKernelGGL((k_cpd_pass2_mx<2, PDX_LDS_LANES, 0>), G, BL, l2b, c->stream, af);
                PDX_HIP(c, hipGetLastError());
                PDX_TRY(pdx_timed_end(c, "step"));
                break;
            }
            const int fc = (c->multi() ? PDX_FEAT_MULTI : 0) | (c->P.zem == PDX_ZEM_WAN ? PDX_FEAT_ZW : 0);
            // pass 2 entries 112 B, 160 B with ZEMWan
            const int e1 = 64 + (c->multi() ? 1 : 0), e2 = c->P.zem == PDX_ZEM_WAN ? 160 : 112;
            PDX_TRY(pdx_timed_begin(c));
            switch (fc) {
            case 0: hipLaunchKernelGGL((k_cpd_lds1<PDX_LDS_LANES, 0>), G, BL, lds_bytes(c, e1, bud), c->stream, af); break;
            case PDX_FEAT_MULTI: hipLaunchKernelGGL((k_cpd_lds1<PDX_LDS_LANES, PDX_FEAT_MULTI>), G, BL, lds_bytes(c, e1, bud), c->stream, af); break;
            case PDX_FEAT_ZW: hipLaunchKernelGGL((k_cpd_lds1<PDX_LDS_LANES, PDX_FEAT_ZW>), G, BL, lds_bytes(c, e1, bud), c->stream, af); break;
            default: hipLaunchKernelGGL((k_cpd_lds1<PDX_LDS_LANES, PDX_FEAT_MULTI | PDX_FEAT_ZW>), G, BL, lds_bytes(c, e1, bud), c->stream, af); break;
            }
            PDX_HIP(c, hipGetLastError());
            PDX_TRY(pdx_timed_end(c, "pass1"));
            PDX_TRY(pdx_timed_begin(c));
            switch (fc) {
            case 0: hipLaunchKernelGGL((k_cpd_lds2_own<PDX_LDS_LANES, 0>), G, BL, lds_bytes(c, e2, bud), c->stream, af); break;
            case PDX_FEAT_MULTI: hipLaunchKernelGGL((k_cpd_lds2_own<PDX_LDS_LANES, PDX_FEAT_MULTI>), G, BL, lds_bytes(c, e2, bud), c->stream, af); break;
            case PDX_FEAT_ZW: hipLaunchKernelGGL((k_cpd_lds2_own<PDX_LDS_LANES, PDX_FEAT_ZW>), G, BL, lds_bytes(c, e2, bud), c->stream, af); break;
            default: hipLaunchKernelGGL((k_cpd_lds2_own<PDX_LDS_LANES, PDX_FEAT_MULTI | PDX_FEAT_ZW>), G, BL, lds_bytes(c, e2, bud), c->stream, af); break;
            }
            PDX_HIP(c, hipGetLastError());
            PDX_TRY(pdx_timed_end(c, "step"));
            break;
        }
        PDX_TRY(pdx_timed_begin(c));
        hipLaunchKernelGGL(k_cpd_pass1<PDX_LDS_LANES>, G, dim3(PDX_BLOCK * PDX_LDS_LANES), 0, c->stream, a);
        PDX_HIP(c, hipGetLastError());
        PDX_TRY(pdx_timed_end(c, "pass1"));
        PDX_TRY(pdx_timed_begin(c));
        hipLaunchKernelGGL(k_cpd_pass2<PDX_LDS_LANES>, G, dim3(PDX_BLOCK * PDX_LDS_LANES), 0, c->stream, a);
        PDX_HIP(c, hipGetLastError());
        PDX_TRY(pdx_timed_end(c, "step"));
        break;
    default: return pdx_fail(c, PDX_ERR_STATE, "no material");
    }
    return PDX_OK;
}

pdx_status pdx_launch_prologue(pdx_ctx *c, int64_t step, double dt, const double *bcv_dev)
{
    (void)step;
    const int64_t nt = 3 * c->n;
    hipLaunchKernelGGL(k_vv_prologue, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0,
                       c->stream, c->n, c->v, c->a, c->vh, c->u, c->rec[c->cur], c->bcslot,
                       c->slot_src, bcv_dev, dt, 0.5 * dt);
    PDX_HIP(c, hipGetLastError());
    return PDX_OK;
}

extern "C" pdx_status pdx_force_density(pdx_ctx *c, double t, double dt)
{
    (void)t;
    if (!c) return PDX_ERR_ARG;
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    PDX_HIP(c, hipSetDevice(c->device));
    PDX_TRY(pdx_reset_flags(c));
    pdx_step_args a;
    PDX_TRY(pdx_fill_step_args(c, &a, 0, PDX_F_FORCE_ONLY, dt, nullptr));
    PDX_TRY(pdx_launch_step(c, a));
    int64_t nan_step = -1;
    PDX_TRY(pdx_check_flags(c, &nan_step));
    if (nan_step >= 0) return pdx_fail(c, PDX_ERR_NAN, "NaN in the force density at t=%g", t);
    return PDX_OK;
}

// the steps of a run, with the boundary-condition value table already on the device
static pdx_status vv_run_steps(pdx_ctx *c, int64_t first, int64_t n_steps, double dt,
                               const double *bcv, int64_t *nan_step_out)
{
    PDX_TRY(pdx_reset_flags(c));
    PDX_TRY(pdx_launch_prologue(c, first, dt, bcv));
    for (int64_t k = 0; k < n_steps; k++) {
        const bool last = (k == n_steps - 1);
        const int flags = last ? PDX_F_STORE : PDX_F_HAS_NEXT;
        const double *next_bc = last ? bcv : bcv + c->n_src * (k + 1);
        // b_ext conditions at t = (first + k) dt, used by this launch's acceleration
        PDX_TRY(pdx_apply_fd_bcs(c, c->n_src > 0 ? bcv + c->n_src * k : nullptr));
        pdx_step_args a;
        PDX_TRY(pdx_fill_step_args(c, &a, first + k, flags, dt, next_bc));
        PDX_TRY(pdx_launch_step(c, a));
        if (!last) c->cur ^= 1;
    }
    int64_t nan_step = -1;
    PDX_TRY(pdx_check_flags(c, &nan_step));
    if (nan_step >= 0) {
        if (nan_step_out) *nan_step_out = nan_step;
        return pdx_fail(c, PDX_ERR_NAN, "NaN in the force density at step %lld",
                        (long long)nan_step);
    }
    return PDX_OK;
}

static pdx_status vv_run_check(pdx_ctx *c, int64_t n_steps, double dt, const double *bc_values)
{
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    if (n_steps <= 0) return PDX_OK;
    if (!(dt > 0)) return pdx_fail(c, PDX_ERR_ARG, "`Δt` of VelocityVerlet smaller than zero!");
    if (c->n_src > 0 && !bc_values) return pdx_fail(c, PDX_ERR_ARG, "missing BC values");
    return PDX_OK;
}

extern "C" pdx_status pdx_vv_run(pdx_ctx *c, int64_t first, int64_t n_steps, double dt,
                                 const double *bc_values, int64_t *nan_step_out)
{
    if (!c) return PDX_ERR_ARG;
    if (nan_step_out) *nan_step_out = -1;
    PDX_TRY(vv_run_check(c, n_steps, dt, bc_values));
    if (n_steps <= 0) return PDX_OK;
    PDX_HIP(c, hipSetDevice(c->device));
    // BC value table: one row of n_src values per step
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
    return vv_run_steps(c, first, n_steps, dt, c->bcvals, nan_step_out);
}

// ----------------------------------------------------------------------------------------
// pdx_vv_run_noalloc: the same run without any device allocation.  pdx_vv_run keeps the
// boundary-condition value table of a run in context memory that grows (hipFree +
// hipMalloc) when a run needs a larger table; this variant stages the table in `scratch`,
// a fixed array of PDX_NOALLOC_BC_VALUES doubles that every GPU holds once for the life
// of the process.  All contexts on one GPU share that GPU's array: calls that overlap on
// one GPU (other host threads, other streams) race on it.  The table is written in full
// by each call before its kernels read it, so nothing needs clearing.
// ----------------------------------------------------------------------------------------
static __device__ __align__(256) double scratch[PDX_NOALLOC_BC_VALUES];

#define PDX_NOALLOC_DEVICES 64
static double *g_scratch_ptr[PDX_NOALLOC_DEVICES];  // device address of `scratch` per GPU
static std::mutex g_scratch_mu;

// The device address of this GPU's `scratch`, looked up once per device with the device
// current.  In host code the array's name is only a host-side stand-in, so every host use
// (kernel arguments, copies) takes this address instead.
static pdx_status noalloc_scratch(pdx_ctx *c, double **out)
{
    if (c->device < 0 || c->device >= PDX_NOALLOC_DEVICES)
        return pdx_fail(c, PDX_ERR_ARG, "device %d beyond the no-alloc table", c->device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[c->device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return pdx_fail(c, PDX_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                            hipGetErrorString(e));
        g_scratch_ptr[c->device] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[c->device];
    return PDX_OK;
}

extern "C" pdx_status pdx_vv_run_noalloc(pdx_ctx *c, int64_t first, int64_t n_steps, double dt,
                                         const double *bc_values, int64_t *nan_step_out)
{
    if (!c) return PDX_ERR_ARG;
    if (nan_step_out) *nan_step_out = -1;
    PDX_TRY(vv_run_check(c, n_steps, dt, bc_values));
    if (n_steps <= 0) return PDX_OK;
    if (c->n_src > PDX_NOALLOC_BC_VALUES / n_steps)
        return pdx_fail(c, PDX_ERR_ARG,
                        "pdx_vv_run_noalloc: %lld steps of %lld boundary-condition values "
                        "exceed the fixed table of %d values; split the run",
                        (long long)n_steps, (long long)c->n_src, PDX_NOALLOC_BC_VALUES);
    PDX_HIP(c, hipSetDevice(c->device));
    double *table = nullptr;
    PDX_TRY(noalloc_scratch(c, &table));
    if (c->n_src > 0)
        PDX_HIP(c, hipMemcpyAsync(table, bc_values, sizeof(double) * c->n_src * n_steps,
                                  hipMemcpyHostToDevice, c->stream));
    return vv_run_steps(c, first, n_steps, dt, table, nan_step_out);
}
