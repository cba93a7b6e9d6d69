// Force density and the fused Velocity Verlet step: BB, OSB and CPD.
//
// One wavefront per point gathers the point's bonds (CSR row of the Morton-ordered bond
// list, 64 bonds per iteration, coalesced), so every point reduces its own bonds without
// atomics.  The scatter of the reference's OSB and CPD force densities (b_int[j] -= ...,
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
#include <cmath>
#include <cstring>

#include "pdx_internal.h"

// ----------------------------------------------------------------------------------------
// per-point tail: NaN flag (check_for_nans, nans.jl:10-18), acceleration, velocity and the
// first half of the next step.  `sub` is the lane index within the point's lane group;
// sub 0..2 handle the x, y, z dofs.
// ----------------------------------------------------------------------------------------
template <bool POINT_STATE>
__device__ __forceinline__ void step_tail(const pdx_step_args &A, int64_t i, int sub,
                                          double4 ri, double bx, double by, double bz, int na,
                                          int64_t nn)
{
    if (POINT_STATE && sub == 0 && (A.flags & (PDX_F_STORE | PDX_F_FORCE_ONLY))) {
        A.nact_out[i] = na;
        // calc_damage!: bond_system.jl:338-343
        A.dmg_out[i] = 1.0 - (double)na / (double)nn;
    }
    if (sub < 3) {
        const double bd = sub == 0 ? bx : (sub == 1 ? by : bz);
        if (isnan(bd)) atomicMin(A.flags_out, (unsigned long long)A.step);
        const int64_t k = 3 * i + sub;
        if (A.flags & PDX_F_FORCE_ONLY) {
            A.bint_out[k] = (A.flags & PDX_F_SCALE_VOL) ? bd * ri.w : bd;
            return;
        }
        const double be = (A.flags & PDX_F_BEXT) ? A.bext[k] : 0.0;
        const double ad = (bd + be) / A.rho;
        const double vd = A.vh[k] + ad * A.dth;
        if (A.flags & PDX_F_STORE) {
            A.v_out[k] = vd;
            A.a_out[k] = ad;
            A.bint_out[k] = bd;
        }
        if (A.flags & PDX_F_HAS_NEXT) {
            double vn = vd + ad * A.dth;
            const int32_t slot = A.bcslot[i];
            if (slot >= 0) {
                const int32_t s = A.slot_src[3 * slot + sub];
                if (s >= 0) {
                    const double val = A.bcv[s];
                    if (isfinite(val)) vn = val;
                }
            }
            const double du = vn * A.dt;
            A.u[k] = A.u[k] + du;
            const double xd = (sub == 0 ? ri.x : (sub == 1 ? ri.y : ri.z)) + du;
            reinterpret_cast<double *>(A.rec_out)[4 * i + sub] = xd;
            A.vh[k] = vn;
        }
    } else if (sub == 3 && (A.flags & PDX_F_HAS_NEXT) && !(A.flags & PDX_F_FORCE_ONLY)) {
        reinterpret_cast<double *>(A.rec_out)[4 * i + 3] = ri.w;
    }
}

// 1/x to about 1 ulp: hardware reciprocal and two Newton steps
__device__ __forceinline__ double pdx_rcp(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// ----------------------------------------------------------------------------------------
// BB: calc_failure! (bond_based.jl:125-144) + calc_damage! + force_density_point!
// (bond_based.jl:146-161) + the fused tail.
//
// PPW points per wavefront (64/PPW lanes per point).  A lane takes the bonds h, h + LPP, ...,
// h + (U-1) LPP of its point per chunk (h = lane index within the point's lanes) and issues
// all their loads and neighbour gathers before any arithmetic.  The operands of the tail
// (u, v_half, b_ext, BC slot) are loaded before the bond loop so that their latency
// overlaps it.
//
// FP64: l = sqrt((dx*dx + dy*dy) + dz*dz) is correctly rounded.  The stretch used in the
// force is (l - L) * (1/L) (within a few ulp of the reference's (l - L)/L); the breakage
// test eps > eps_c uses the correctly rounded quotient whenever the two could disagree
// (|eps - eps_c| <= 1e-11 |eps_c|), so bonds break exactly when the reference's break.  The
// force term is ((w bc) eps V_j) * (1/l), accumulated with dx by fused multiply-adds (the
// reference divides each component by l and adds).  An inactive bond has w = 0, as the
// reference's Bool * factor, so NaN positions still show.
// MIXED: the stretch comes from FP64 differences du = (x_j - x_i) - (X_j - X_i), then
// eps = (2 dX.du + du.du) / (L (l + L)) in FP32 and the force in FP32, summed in FP64.
// ----------------------------------------------------------------------------------------
__device__ __forceinline__ double pdx_wave_total(double x)
{
    x = pdx_row_sum(x);
    return (pdx_readlane(x, 0) + pdx_readlane(x, 16)) + (pdx_readlane(x, 32) + pdx_readlane(x, 48));
}

__device__ __forceinline__ int pdx_wave_total_i(int x)
{
    x = pdx_row_sum_i(x);
    return (__builtin_amdgcn_readlane(x, 0) + __builtin_amdgcn_readlane(x, 16)) +
           (__builtin_amdgcn_readlane(x, 32) + __builtin_amdgcn_readlane(x, 48));
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

// One bond of the BB force density (see k_bb_step for the arithmetic contract).
template <bool MIXED>
__device__ __forceinline__ void bb_bond(const pdx_step_args &A, int64_t b, uint32_t wd, double Ld,
                                        float Lfl, double4 rj, double4 Xj, double4 ri,
                                        double4 Xi, double &bx, double &by, double &bz, int &na)
{
    const double dx = rj.x - ri.x, dy = rj.y - ri.y, dz = rj.z - ri.z;
    bool act = (wd & PDX_BW_ACT) != 0;
    if (!MIXED) {
        const double d2 = (dx * dx + dy * dy) + dz * dz;
        double rl;
        const double l = pdx_sqrt_rsq(d2, rl);
        const double e = l - Ld;
        const double eps = e * pdx_rcp1(Ld);
        // breakage: eps > eps_c with the correctly rounded quotient wherever it could differ
        if (act && (wd & PDX_BW_FP) && eps >= A.epsc_lo) {
            const bool brk = (eps > A.epsc_hi) ? true : ((e / Ld) > A.epsc);
            if (brk) {
                act = false;
                A.bw[b] = wd & ~PDX_BW_ACT;
            }
        }
        const double wbc = act ? A.bc : 0.0;  // (w * bc) with w = bond_active
        const double s = ((wbc * eps) * rj.w) * rl;
        bx = __builtin_fma(s, dx, bx);
        by = __builtin_fma(s, dy, by);
        bz = __builtin_fma(s, dz, bz);
    } else {
        const double DXd = Xj.x - Xi.x, DYd = Xj.y - Xi.y, DZd = Xj.z - Xi.z;
        const float ux = (float)(dx - DXd), uy = (float)(dy - DYd), uz = (float)(dz - DZd);
        const float DX = (float)DXd, DY = (float)DYd, DZ = (float)DZd;
        const float num = 2.0f * ((DX * ux + DY * uy) + DZ * uz) + ((ux * ux + uy * uy) + uz * uz);
        const float l = __fsqrt_rn(Lfl * Lfl + num);
        const float eps = __fdividef(num, Lfl * (l + Lfl));
        if (act && (wd & PDX_BW_FP) && eps > (float)A.epsc) {
            act = false;
            A.bw[b] = wd & ~PDX_BW_ACT;
        }
        const float wbc = act ? (float)A.bc : 0.0f;
        const float sf = __fdividef((wbc * eps) * (float)rj.w, l);
        bx += (double)(sf * (float)dx);
        by += (double)(sf * (float)dy);
        bz += (double)(sf * (float)dz);
    }
    na += act ? 1 : 0;
}

template <bool MIXED, int PPW, int U>
__global__ __launch_bounds__(PDX_BLOCK) void k_bb_step(pdx_step_args A)
{
    constexpr int LPP = 64 / PPW;
    const int lane = threadIdx.x & 63;
    const int grp = (PPW == 2) ? (lane >> 5) : 0;
    const int h = lane & (LPP - 1);
    const int64_t w = pdx_wave_point();
    if ((int64_t)PPW * w >= A.n) return;
    const int64_t i = PPW * w + grp;
    const bool valid = i < A.n;
    const int flags = A.flags;
    int64_t b0 = 0, b1 = 0;
    double4 ri = make_double4(0, 0, 0, 0), Xi = make_double4(0, 0, 0, 0);
    double u_d = 0.0, vh_d = 0.0, be_d = 0.0;
    int32_t slot = -1;
    if (valid) {
        b0 = A.row_ptr[i];
        b1 = A.row_ptr[i + 1];
        ri = A.rec_in[i];
        if (MIXED) Xi = A.Xr[i];
        if (h < 3 && !(flags & PDX_F_FORCE_ONLY)) {
            const int64_t k = 3 * i + h;
            vh_d = A.vh[k];
            if (flags & PDX_F_HAS_NEXT) {
                u_d = A.u[k];
                slot = A.bcslot[i];
            }
            if (flags & PDX_F_BEXT) be_d = A.bext[k];
        }
    }
    double bx = 0.0, by = 0.0, bz = 0.0;
    int na = 0;
    for (int64_t base = b0 + h; base < b1; base += LPP * U) {
        uint32_t wd[U];
        double Ld[U];
        float Lfl[U];
        double4 rj[U], Xj[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            Ld[u] = 0.0;
            Lfl[u] = 0.0f;
            Xj[u] = make_double4(0, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t b = base + LPP * u;
            wd[u] = 0;
            if (b < b1) {
                wd[u] = A.bw[b];
                if (MIXED) Lfl[u] = A.Lf[b];
                else Ld[u] = A.L[b];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (base + LPP * u < b1) {
                const uint32_t j = wd[u] & PDX_BW_NBR;
                rj[u] = A.rec_in[j];
                if (MIXED) Xj[u] = A.Xr[j];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t b = base + LPP * u;
            if (b < b1)
                bb_bond<MIXED>(A, b, wd[u], MIXED ? 0.0 : Ld[u], MIXED ? Lfl[u] : 0.0f, rj[u],
                               MIXED ? Xj[u] : make_double4(0, 0, 0, 0), ri, Xi, bx, by, bz, na);
        }
    }
    if (PPW == 2) {
        bx = pdx_half_sum(bx, grp);
        by = pdx_half_sum(by, grp);
        bz = pdx_half_sum(bz, grp);
        na = pdx_half_sum_i(na, grp);
    } else {
        bx = pdx_wave_total(bx);
        by = pdx_wave_total(by);
        bz = pdx_wave_total(bz);
        na = pdx_wave_total_i(na);
    }
    if (!valid) return;
    // tail: calc_damage!, check_for_nans, _update_acc_and_vel!, then the first half of the
    // next step (update_vel_half!, apply_boundary_conditions!, update_disp_and_pos!)
    if (h == 0 && (flags & (PDX_F_STORE | PDX_F_FORCE_ONLY))) {
        A.nact_out[i] = na;
        A.dmg_out[i] = 1.0 - (double)na / (double)(b1 - b0);
    }
    if (h < 3) {
        const double bd = h == 0 ? bx : (h == 1 ? by : bz);
        if (isnan(bd)) atomicMin(A.flags_out, (unsigned long long)A.step);
        const int64_t k = 3 * i + h;
        if (flags & PDX_F_FORCE_ONLY) {
            A.bint_out[k] = (flags & PDX_F_SCALE_VOL) ? bd * ri.w : bd;
            return;
        }
        const double ad = (bd + be_d) / A.rho;
        const double vd = vh_d + ad * A.dth;
        if (flags & PDX_F_STORE) {
            A.v_out[k] = vd;
            A.a_out[k] = ad;
            A.bint_out[k] = bd;
        }
        if (flags & PDX_F_HAS_NEXT) {
            double vn = vd + ad * A.dth;
            if (slot >= 0) {
                const int32_t src = A.slot_src[3 * slot + h];
                if (src >= 0) {
                    const double val = A.bcv[src];
                    if (isfinite(val)) vn = val;
                }
            }
            const double du = vn * A.dt;
            A.u[k] = u_d + du;
            const double xd = (h == 0 ? ri.x : (h == 1 ? ri.y : ri.z)) + du;
            reinterpret_cast<double *>(A.rec_out)[4 * i + h] = xd;
            A.vh[k] = vn;
        }
    } else if (h == 3 && (flags & PDX_F_HAS_NEXT) && !(flags & PDX_F_FORCE_ONLY)) {
        reinterpret_cast<double *>(A.rec_out)[4 * i + 3] = ri.w;
    }
}

// ----------------------------------------------------------------------------------------
// OSB pass 1: calc_failure! (ordinary_state_based.jl:107-126), calc_damage!,
// calc_weighted_volume (:176-189), calc_dilatation (:191-204) -> c1, c2 (:135-136).
// |dX|^2 is taken as L*L (the reference forms dot(dX, dX); they differ by <= 2 ulp).
// ----------------------------------------------------------------------------------------
__global__ __launch_bounds__(PDX_BLOCK) void k_osb_pass1(pdx_step_args A)
{
    const int64_t i = pdx_wave_point();
    if (i >= A.n) return;
    const int lane = threadIdx.x & 63;
    const double4 ri = A.rec_in[i];
    const int64_t b0 = A.row_ptr[i], b1 = A.row_ptr[i + 1];
    double wvol = 0.0, S = 0.0;
    int na = 0;
    for (int64_t b = b0 + lane; b < b1; b += 64) {
        const uint32_t wd = A.bw[b];
        const uint32_t j = wd & PDX_BW_NBR;
        const double4 rj = A.rec_in[j];
        const double dx = rj.x - ri.x, dy = rj.y - ri.y, dz = rj.z - ri.z;
        const double Lb = A.L[b];
        const double l = sqrt((dx * dx + dy * dy) + dz * dz);
        const double eps = (l - Lb) / Lb;
        bool act = (wd & PDX_BW_ACT) != 0;
        if (act && (wd & PDX_BW_FP) && eps > A.epsc) {
            act = false;
            A.bw[b] = wd & ~PDX_BW_ACT;
        }
        if (act) {
            const double w = A.delta / Lb;  // linear_kernel
            wvol += (w * (Lb * Lb)) * rj.w;
            S += ((w * Lb) * (l - Lb)) * rj.w;
            na++;
        }
    }
    wvol = pdx_wave_sum(wvol);
    S = pdx_wave_sum(S);
    na = pdx_wave_sum_i(na);
    if (lane == 0) {
        double c1 = 0.0, c2 = 0.0;
        if (wvol != 0.0) {
            const double dil = (3.0 / wvol) * S;
            c1 = (15.0 * A.G) / wvol;
            c2 = dil * ((3.0 * A.K) / wvol - c1 / 3.0);
        }
        A.pstate[2 * i] = c1;
        A.pstate[2 * i + 1] = c2;
        if (A.flags & (PDX_F_STORE | PDX_F_FORCE_ONLY)) {
            A.nact_out[i] = na;
            A.dmg_out[i] = 1.0 - (double)na / (double)(b1 - b0);
        }
    }
}

// OSB pass 2: force_density_point! (ordinary_state_based.jl:128-149) as a gather:
// b_i = sum_j (w/l) dx_ij [ (c2_i L + c1_i (l-L)) V_j + (c2_j L + c1_j (l-L)) V_i ]
__global__ __launch_bounds__(PDX_BLOCK) void k_osb_pass2(pdx_step_args A)
{
    const int64_t i = pdx_wave_point();
    if (i >= A.n) return;
    const int lane = threadIdx.x & 63;
    const double4 ri = A.rec_in[i];
    const double c1i = A.pstate[2 * i], c2i = A.pstate[2 * i + 1];
    const int64_t b0 = A.row_ptr[i], b1 = A.row_ptr[i + 1];
    double bx = 0.0, by = 0.0, bz = 0.0;
    for (int64_t b = b0 + lane; b < b1; b += 64) {
        const uint32_t wd = A.bw[b];
        if (!(wd & PDX_BW_ACT)) continue;
        const uint32_t j = wd & PDX_BW_NBR;
        const double4 rj = A.rec_in[j];
        const double2 cj = reinterpret_cast<const double2 *>(A.pstate)[j];
        const double dx = rj.x - ri.x, dy = rj.y - ri.y, dz = rj.z - ri.z;
        const double Lb = A.L[b];
        const double l = sqrt((dx * dx + dy * dy) + dz * dz);
        const double w = A.delta / Lb;
        const double e = l - Lb;
        const double fi = c2i * Lb + c1i * e;
        const double fj = cj.y * Lb + cj.x * e;
        const double s = (w / l) * (fi * rj.w + fj * ri.w);
        bx += s * dx;
        by += s * dy;
        bz += s * dz;
    }
    bx = pdx_wave_sum(bx);
    by = pdx_wave_sum(by);
    bz = pdx_wave_sum(bz);
    // damage and n_active were written by pass 1
    step_tail<false>(A, i, lane, ri, bx, by, bz, 0, b1 - b0);
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
__device__ void pk1(int model, const double *F, double lambda, double mu, double G, double K,
                    double *P)
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

// CPD pass 1: calc_failure! (bond_system.jl:309-327), calc_damage!,
// calc_deformation_gradient! (correspondence.jl:193-214), too_much_damage! (:288-298),
// calc_first_piola_kirchhoff! (:216-224).  Stores F and P*K^-1 of every point.
__global__ __launch_bounds__(PDX_BLOCK) void k_cpd_pass1(pdx_step_args A)
{
    const int64_t i = pdx_wave_point();
    if (i >= A.n) return;
    const int lane = threadIdx.x & 63;
    const double4 ri = A.rec_in[i];
    const double4 Xi = A.Xr[i];
    const bool dead = A.killed[i] != 0;
    const int64_t b0 = A.row_ptr[i], b1 = A.row_ptr[i + 1];
    double Km[6] = {0, 0, 0, 0, 0, 0};  // 00 10 20 11 21 22
    double Ft[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int na = 0;
    for (int64_t b = b0 + lane; b < b1; b += 64) {
        const uint32_t wd = A.bw[b];
        const uint32_t j = wd & PDX_BW_NBR;
        const double4 rj = A.rec_in[j];
        const double4 Xj = A.Xr[j];
        const double dx[3] = {rj.x - ri.x, rj.y - ri.y, rj.z - ri.z};
        const double DX[3] = {Xj.x - Xi.x, Xj.y - Xi.y, Xj.z - Xi.z};
        const double Lb = A.L[b];
        const double l = sqrt((dx[0] * dx[0] + dx[1] * dx[1]) + dx[2] * dx[2]);
        const double eps = (l - Lb) / Lb;
        bool act = (wd & PDX_BW_ACT) != 0;
        if (act && (wd & PDX_BW_FP) && eps > A.epsc) {
            act = false;
            A.bw[b] = wd & ~PDX_BW_ACT;
        }
        if (act && !dead) {
            na++;
            const double temp = (A.delta / Lb) * rj.w;
            Km[0] += temp * (DX[0] * DX[0]);
            Km[1] += temp * (DX[1] * DX[0]);
            Km[2] += temp * (DX[2] * DX[0]);
            Km[3] += temp * (DX[1] * DX[1]);
            Km[4] += temp * (DX[2] * DX[1]);
            Km[5] += temp * (DX[2] * DX[2]);
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < 3; r++) Ft[r + 3 * c] += temp * (dx[r] * DX[c]);
        }
    }
#pragma unroll
    for (int k = 0; k < 6; k++) Km[k] = pdx_wave_sum(Km[k]);
#pragma unroll
    for (int k = 0; k < 9; k++) Ft[k] = pdx_wave_sum(Ft[k]);
    na = pdx_wave_sum_i(na);
    const double K9[9] = {Km[0], Km[1], Km[2], Km[1], Km[3], Km[4], Km[2], Km[4], Km[5]};
    double Kinv[9], F[9];
    m3_inv(K9, Kinv);
    m3_mul(Ft, Kinv, F);
    const double dmg = 1.0 - (double)na / (double)(b1 - b0);
    bool hasnan = false;
#pragma unroll
    for (int k = 0; k < 9; k++) hasnan |= isnan(F[k]);
    const bool kill = dead || dmg > A.maxdmg || hasnan;
    const bool store = (A.flags & (PDX_F_STORE | PDX_F_FORCE_ONLY)) != 0;
    double PK[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (!kill) {
        double P[9];
        pk1(A.model, F, A.lambda, A.mu, A.G, A.K, P);
        m3_mul(P, Kinv, PK);
        if (store) {
            // cauchy_stress: stress.jl:112-116, sigma = (1/J) P F'
            const double s = 1.0 / m3_det(F);
            if (lane < 9) {
                const int r = lane % 3, c = lane / 3;
                A.cauchy[9 * i + lane] =
                    ((s * P[r]) * F[c] + (s * P[r + 3]) * F[c + 3]) + (s * P[r + 6]) * F[c + 6];
            }
        }
    }
    if (lane < 9) {
        A.pstate[18 * i + lane] = F[lane];
        A.pstate[18 * i + 9 + lane] = PK[lane];
        if (store) A.defgrad[9 * i + lane] = F[lane];
    }
    if (lane == 0) {
        if (kill && !dead) A.killed[i] = 1;
        if (store) {
            A.nact_out[i] = kill ? 0 : na;
            A.dmg_out[i] = dmg;
        }
    }
}

// CPD pass 2: c_force_density! with ZEMSilling (correspondence.jl:226-251) as a gather:
// own term  t_ij V_j           with t_ij = (w PKinv_i) dX + (w beta) (dx - F_i dX)
// reverse  -t_ji V_i  = V_i [ (w PKinv_j) dX + (w beta) (dx - F_j dX) ]
__global__ __launch_bounds__(PDX_BLOCK) void k_cpd_pass2(pdx_step_args A)
{
    const int64_t i = pdx_wave_point();
    if (i >= A.n) return;
    const int lane = threadIdx.x & 63;
    const double4 ri = A.rec_in[i];
    const double4 Xi = A.Xr[i];
    const bool dead_i = A.killed[i] != 0;
    double Fi[9], PKi[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        Fi[k] = A.pstate[18 * i + k];
        PKi[k] = A.pstate[18 * i + 9 + k];
    }
    const int64_t b0 = A.row_ptr[i], b1 = A.row_ptr[i + 1];
    double b3[3] = {0.0, 0.0, 0.0};
    for (int64_t b = b0 + lane; b < b1; b += 64) {
        const uint32_t wd = A.bw[b];
        if (!(wd & PDX_BW_ACT)) continue;
        const uint32_t j = wd & PDX_BW_NBR;
        const double4 rj = A.rec_in[j];
        const double4 Xj = A.Xr[j];
        const double dx[3] = {rj.x - ri.x, rj.y - ri.y, rj.z - ri.z};
        const double DX[3] = {Xj.x - Xi.x, Xj.y - Xi.y, Xj.z - Xi.z};
        const double w = A.delta / A.L[b];
        const double wb = w * A.beta_zem;
        if (!dead_i) {
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const double FX = (Fi[r] * DX[0] + Fi[r + 3] * DX[1]) + Fi[r + 6] * DX[2];
                const double tz = wb * (dx[r] - FX);
                const double t = (((w * PKi[r]) * DX[0] + (w * PKi[r + 3]) * DX[1]) +
                                  (w * PKi[r + 6]) * DX[2]) + tz;
                b3[r] += t * rj.w;
            }
        }
        if (!A.killed[j]) {
            const double *pj = A.pstate + 18 * (int64_t)j;
            double Fj[9], PKj[9];
#pragma unroll
            for (int k = 0; k < 9; k++) {
                Fj[k] = pj[k];
                PKj[k] = pj[9 + k];
            }
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const double FX = (Fj[r] * DX[0] + Fj[r + 3] * DX[1]) + Fj[r + 6] * DX[2];
                const double tz = wb * (dx[r] - FX);
                const double t = (((w * PKj[r]) * DX[0] + (w * PKj[r + 3]) * DX[1]) +
                                  (w * PKj[r + 6]) * DX[2]) + tz;
                b3[r] += t * ri.w;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 3; r++) b3[r] = pdx_wave_sum(b3[r]);
    step_tail<false>(A, i, lane, ri, b3[0], b3[1], b3[2], 0, b1 - b0);
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
    a->row_ptr = c->row_ptr;
    a->bw = c->bw;
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
    a->beta_zem = (c->P.zem_cs * c->P.bc) / c->P.delta;  // correspondence.jl:232
    a->maxdmg = c->P.maxdmg;
    a->model = c->P.model;
    a->pstate = c->pstate;
    a->killed = c->killed;
    a->defgrad = c->defgrad;
    a->cauchy = c->cauchy;
    return PDX_OK;
}

pdx_status pdx_launch_step(pdx_ctx *c, const pdx_step_args &a)
{
    const int64_t grid = pdx_wave_grid(c->n);
    const dim3 G((unsigned)grid), B(PDX_BLOCK);
    const dim3 G2((unsigned)pdx_wave_grid((c->n + 1) / 2));
    switch (c->kind) {
    case PDX_BB: {
        PDX_TRY(pdx_timed_begin(c));
        const int v = c->bb_variant;
        const dim3 G1((unsigned)pdx_wave_grid(c->n));
        if (c->precision == PDX_MIXED) {
            if (v == 22) hipLaunchKernelGGL((k_bb_step<true, 2, 2>), G2, B, 0, c->stream, a);
            else if (v == 11) hipLaunchKernelGGL((k_bb_step<true, 1, 1>), G1, B, 0, c->stream, a);
            else hipLaunchKernelGGL((k_bb_step<true, 1, 2>), G1, B, 0, c->stream, a);
        } else {
            if (v == 22) hipLaunchKernelGGL((k_bb_step<false, 2, 2>), G2, B, 0, c->stream, a);
            else if (v == 11) hipLaunchKernelGGL((k_bb_step<false, 1, 1>), G1, B, 0, c->stream, a);
            else hipLaunchKernelGGL((k_bb_step<false, 1, 2>), G1, B, 0, c->stream, a);
        }
        PDX_HIP(c, hipGetLastError());
        PDX_TRY(pdx_timed_end(c, "step"));
        break;
    }
    case PDX_OSB:
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
        PDX_TRY(pdx_timed_begin(c));
        hipLaunchKernelGGL(k_cpd_pass1, G, B, 0, c->stream, a);
        PDX_HIP(c, hipGetLastError());
        PDX_TRY(pdx_timed_end(c, "pass1"));
        PDX_TRY(pdx_timed_begin(c));
        hipLaunchKernelGGL(k_cpd_pass2, G, B, 0, c->stream, a);
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

extern "C" pdx_status pdx_vv_run(pdx_ctx *c, int64_t first, int64_t n_steps, double dt,
                                 const double *bc_values, int64_t *nan_step_out)
{
    if (!c) return PDX_ERR_ARG;
    if (nan_step_out) *nan_step_out = -1;
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    if (n_steps <= 0) return PDX_OK;
    if (!(dt > 0)) return pdx_fail(c, PDX_ERR_ARG, "`Δt` of VelocityVerlet smaller than zero!");
    if (c->n_src > 0 && !bc_values) return pdx_fail(c, PDX_ERR_ARG, "missing BC values");
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
    PDX_TRY(pdx_reset_flags(c));
    PDX_TRY(pdx_launch_prologue(c, first, dt, c->bcvals));
    for (int64_t k = 0; k < n_steps; k++) {
        const bool last = (k == n_steps - 1);
        const int flags = last ? PDX_F_STORE : PDX_F_HAS_NEXT;
        const double *next_bc = last ? c->bcvals : c->bcvals + c->n_src * (k + 1);
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
